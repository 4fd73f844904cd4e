/*
 * mhd_oracle.h — CPU restatement of the reference RK3 timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product (libmhd3d.so) never links or calls it.
 *
 * It restates, sweep by sweep and in the same operation order, the reference
 * files:
 *   3dmhdsub.h:31-153   step(): TSTEP/CFL, three Wray RK3 stages
 *   3dmhdsub.h:155-192  step_block_ZETA0 / ZETA / GAM
 *   3dmhdsub.h:194-696  fluxes() incl. the off-default branches iD 1/2, lRem,
 *                       lShr, iTCon 2 and the RLAX relaxation (:322-442) of a
 *                       reference run on relax_nPEz z ranks
 *   3dmhdsub.h:698-718  horizontal_mean (one y rank)
 *   3dmhdsub.h:720-873  BCON()
 *   3dmhdutils.h:45-239 operator templates and reductions
 *   3dmhdmpi.h:32-143   ghost refill, single domain (x, y periodic)
 *   3dmhdphsic.h:230-355,523-547,549-763  bsstep/mmid/rzextr/derivs, RAN2, Static
 *   3dmhdphsic.h:357-522  the lRem background (ODEINT/RKQC/RK4/derivs1/kappa)
 * The arrays are the reference's [nz][ny][nx] with a 1-cell ghost layer.
 * Domain = one z slab holding both walls: the one-sided wall stencils of both
 * walls apply (the reference's nPEz>=2 behaviour, 3dmhdsub.h:211-226,572-584).
 *
 * What pins each function: the published iteration-9 answers
 * (tests/test_oracle_known_answers.py), closed forms on uniform and
 * stretched grids (tests/test_oracle_math.py, test_oracle_stretched.py,
 * test_oracle_ext.py, test_oracle_relax.py), and planted mistakes those
 * checks must catch (tests/test_oracle_mutants.py, the ORC_MUTANT switches
 * of mhd_oracle.c).  The lMag block is pinned by the continuum identities
 * its fluxes must satisfy, not to the last bit: no published answer covers
 * it, so the order of its terms within one output follows the reference's
 * source by reading alone.
 */
#ifndef MHD_ORACLE_H_
#define MHD_ORACLE_H_

#include "../include/mhd3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const mhd3d_desc* d);
void orc_destroy(orc_ctx* c);
void orc_set_threads(int n);

/* Named arrays of the reference (3dmhddata.h:12-19): "RU","RV","RW","RO",
 * "TT","BX","BY","BZ","UU","VV","WW","FU","FV","FW","FR","FT","WW1","WW2",
 * "WW3","ZRU",...,"ZBZ".  Full [nz][ny][nx] copies. Returns 0 or -1. */
int orc_set(orc_ctx* c, const char* name, const double* a);
int orc_get(orc_ctx* c, const char* name, double* a);

void orc_communicate(orc_ctx* c);      /* 3dmhdmpi.h:11-143, single domain      */
void orc_fluxes(orc_ctx* c);           /* 3dmhdsub.h:194-696                     */
void orc_bcon(orc_ctx* c);             /* 3dmhdsub.h:720-873                     */
/* TSTEP (3dmhdsub.h:36-106): sets UU, WW1..3 (, VV) like the reference,
 * writes wmin[0..3], umach; returns DT. */
double orc_cfl(orc_ctx* c, double wmin[4], double* umach);
/* One stage k=1..3 of step() (3dmhdsub.h:116-147) with the given DT. */
void orc_stage(orc_ctx* c, int k, double dt);
/* step() (3dmhdsub.h:31-153); fixed_dt!=0 skips TSTEP and uses dt. Returns DT. */
double orc_step(orc_ctx* c, int fixed_dt, double dt);
void orc_clock(orc_ctx* c, long long* nIt, double* timt, double* timc);
/* The globals DT and TIMT that fluxes() reads (for calls outside step()). */
void orc_set_time(orc_ctx* c, double dt, double timt);
/* RLAX state (3dmhdsub.h:195-200): out = {TFAC, TTM_MAX, TTM_MIN}, ICALL,
 * the exit point reached (0 none, 1 "yk not implement" :427-439, 2 "Relax:
 * check boundary temperature" :344,416) and, when non-null, copies of the
 * last CORRECT, FCONM and FRADM profiles (nz each).  Returns 0 or -1. */
int orc_relax_state(orc_ctx* c, double out[3], long long* icall, int* exited, double* correct,
                    double* fconm, double* fradm);

/* RAN2 (3dmhdphsic.h:523-547).  state = {IDUM, IIY, IIR[97]} (99 ints). */
double orc_ran2(int* state);
/* Bulirsch-Stoer polytrope of Static() (3dmhdphsic.h:596-609) on the uniform
 * grid (nGrid=0, PZP=0): T[k], R[k] for k=0..nPz-1.  Returns 0 or -1. */
int orc_static_profile(int nPz, double THETA, double GRAV, double zMax,
                       double* T, double* R);
/* RND planes of Static() (3dmhdphsic.h:705-744) for an nPEy x nPEz rank grid,
 * written into the single-domain array `out` [nzg][nyg][nxg] (interior only):
 * out = 1 + AMPT*(RAN2-0.5), drawn in the reference's rank order. */
void orc_static_rnd(int nPx, int nPy, int nRz_total, int nPEy, int nPEz,
                    double AMPT, double* out);
/* M. Rempel's background of Static() under lRem (3dmhdphsic.h:357-522,
 * 565-592) on the uniform global grid: T[k], R[k], RK[k] = kappa*8.07f*THETA*
 * REPR, DRK[k] and walls = {TU, DZTU, DZTB}.  Returns 0 or -1. */
int orc_rempel_profile(int nPz, double THETA, double REPR, double zMax, double* T, double* R,
                       double* RK, double* DRK, double walls[3]);

#ifdef __cplusplus
}
#endif
#endif
