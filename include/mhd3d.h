/*
 * mhd3d.h — C ABI of the MI355X (gfx950) RK3 timestep for 3-D compressible
 * (magneto)hydrodynamics on a z-stretched grid.
 *
 * The reference program (thu-scc/Rast-3dmhd-cpp) keeps its state in global
 * arrays `double RU[nz][ny][nx], ...` (3dmhddata.h:8,12-19) and advances it
 * with `void step()` (3dmhdsub.h:31), which calls `fluxes()` (3dmhdsub.h:194),
 * `step_block_ZETA0/ZETA/GAM` (3dmhdsub.h:155-192), `BCON()` (3dmhdsub.h:720)
 * and `communicate()` (3dmhdmpi.h:11).  This header is the drop-in boundary for
 * that call surface: the caller hands over the same [nz][ny][nx] x-contiguous
 * FP64 arrays (1-cell ghost layer on every side) and the same scalars that
 * `set()` (3dmhd.cpp:353-495) leaves in the globals, and gets the same arrays
 * back.  Every entry point returns an int status (MHD3D_OK = 0).
 *
 * One context = one domain on one GPU with its own HIP stream.  A context is
 * not thread-safe; independent contexts may be used from different threads.
 */
#ifndef MHD3D_H_
#define MHD3D_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field slots, in the order of 3dmhddata.h:12-16 (RU,RV,RW,RO,TT) and
 * 3dmhddata.h:17 (BX,BY,BZ).  Flux slots use the same numbering:
 * FU,FV,FW,FR,FT (3dmhddata.h:14) and WW1,WW2,WW3 = dB/dt (3dmhdsub.h:548-694). */
enum {
    MHD3D_RU = 0, MHD3D_RV = 1, MHD3D_RW = 2, MHD3D_RO = 3, MHD3D_TT = 4,
    MHD3D_BX = 5, MHD3D_BY = 6, MHD3D_BZ = 7, MHD3D_NFIELD = 8
};

/* Status codes. */
enum {
    MHD3D_OK = 0,
    MHD3D_EINVAL = 1,       /* bad pointer, size or descriptor value            */
    MHD3D_EUNSUPPORTED = 2, /* flag combination this build does not implement   */
    MHD3D_EHIP = 3,         /* a HIP runtime call failed (see mhd3d_last_error) */
    MHD3D_ENOMEM = 4,       /* device allocation failed                         */
    MHD3D_ENONFINITE = 5,   /* state holds NaN/Inf after the step (the DT guard
                               of 3dmhd.cpp:292 cannot see it: minVal ignores NaN) */
    MHD3D_EHALTED = 6,      /* the DT guard stopped the run (mhd3d_set_dt_guard) */
    MHD3D_EEXIT = 7         /* the reference's program exits at this point of the
                               run (exit_err in the RLAX relaxation, 3dmhdsub.h:344,
                               416,427-439); the context stops (see mhd3d_step)   */
};

/* DT guard threshold of the reference's driver: DT < 1.0e-8f (3dmhd.cpp:292). */
#define MHD3D_REFERENCE_DT_MIN ((double)1.0e-8f)

/* Arithmetic form of the stencil evaluation. */
enum {
    MHD3D_FP64 = 0,   /* plain FP64, same operation order as the reference   */
    MHD3D_MIXED = 1   /* FP32 stencils from FP32 copies, FP64 RK registers,
                         FP64 CFL and diagnostics; with an off-default branch
                         the fluxes are FP64 from the same FP32 copies      */
};

/*
 * Run descriptor.  Sizes follow 3dmhdparam.h:25-47 for ONE domain:
 *   nPx, nPy   interior points in x and y (x, y periodic),
 *   nPz        z planes owned by the domain including both walls,
 * so each array is [nPz+2][nPy+2][nPx+2] doubles (x fastest).
 * Flags are the compile-time switches of 3dmhdparam.h:8-17,86-118, here at
 * run time.  Scalars are the doubles `set()` computes (3dmhd.cpp:353-495):
 * the caller passes them, so float-literal roundings (C13 = 1.0f/3.0f, ...)
 * stay exactly those of the host program.
 */
typedef struct mhd3d_desc {
    int32_t nPx, nPy, nPz;
    /* the switches lRot, lMag, lPot, lRem, lShr (3dmhdparam.h:8-17), ixCon,
     * iyCon, izCon, iTCon, iBCon (:86-107) and iD (:118); the flag_ prefix
     * keeps the names apart from the reference's macros of the same names,
     * so this header compiles inside the reference's translation unit */
    int32_t flag_lRot, flag_lMag, flag_lPot, flag_lRem, flag_lShr;
    int32_t flag_ixCon, flag_iyCon, flag_izCon, flag_iTCon, flag_iBCon;
    int32_t flag_iD;
    int32_t precision;                           /* MHD3D_FP64 / MHD3D_MIXED  */
    /* RLAX > 0 only: the z ranks nPEz of the reference run (3dmhdparam.h:21)
     * whose relaxation arithmetic is reproduced -- the scan of CORRECT runs
     * per rank and meets at the rank boundaries (3dmhdsub.h:354-421); >= 2,
     * dividing nPz.  Ignored when RLAX == 0. */
    int32_t relax_nPEz;
    /* GRID / RUNGKU / SFF (3dmhd.cpp:354-369, 3dmhdparam.h:111) */
    double C13, C23, C43;
    double GAM1, GAM2, GAM3, ZETA1, ZETA2;
    double SFF;
    double hx, h2x, hy, h2y, hz, h2z;
    double ORZF;        /* (double)(1.0f/float(nPz_global-1)): DZ in BCON, 3dmhdsub.h:797,849 */
    /* CPAR / CROT / CMAG / CPEN / CPER (3dmhd.cpp:144-145,401-459) */
    double RE, ORE, REPR, CV, OCV, GAMMA, THETA, GRAV;
    double OMX, OMZ;
    double RM, ORM, OBETA;
    double TP, TB, TC;
    /* heat-flux relaxation (3dmhdsub.h:322-442): RLAX = par[59], TSTART =
     * par[57], TOFF = par[58] (3dmhd.cpp:428-430); with ZEE and
     * relax_nPEz.  The reference exits in the first fluxes() call of a step
     * that starts within DT after a multiple of 25 in TIMT (:427-439), so a
     * run can only use it from a restart; MHD3D_EEXIT reports that point. */
    double RLAX, TSTART, TOFF;
    double TIMI;        /* initial simulation time (3dmhd.cpp:114)                   */
    /* AJACOBI (3dmhddata.h:20-22): lengths nPx+2, nPy+2, nPz+2 */
    const double *dxxdx, *d2xxdx2, *DDX;
    const double *dyydy, *d2yydy2, *DDY;
    const double *dzzdz, *d2zzdz2, *DDZ;
    /* CPEN (3dmhddata.h:32): length nPz+2 */
    const double *RKAPA, *DKAPA;
    /* Off-default branches; read only when their flag asks for them.
     * iD != 0 (density diffusion, 3dmhdsub.h:229-264): DH = par[55],
     * DV = par[56] (3dmhd.cpp:449); iD == 2 also zMax = par[18] and ZEE. */
    double DH, DV, zMax;
    /* lRem wall temperatures that Static() leaves in TU, DZTU, DZTB
     * (3dmhdphsic.h:590-592), used by BCON (3dmhdsub.h:783-858). */
    double TU, DZTU, DZTB;
    /* iTCon 2 heat loss (3dmhdsub.h:443-461) and the TP != 0 plume of iTCon
     * 0/1 (:744-776): XP, YP, ZP = par[13..15], QFH = par[52], HH = par[54];
     * ORZD = 1.0f/double(nPz-1) with the global nPz (:768). */
    double XP, YP, ZP, QFH, HH, ORZD;
    /* grid coordinates (3dmhddata.h:20-22): lengths nPx+2, nPy+2, nPz+2;
     * needed by iD 2 (ZEE), iTCon 2 (all) and the plume (EXX, WYY). */
    const double *EXX, *WYY, *ZEE;
} mhd3d_desc;

typedef struct mhd3d_ctx mhd3d_ctx;

/* CFL record of one TSTEP evaluation (3dmhdsub.h:36-106). */
typedef struct mhd3d_cfl_info {
    double wmin[4];   /* minVal(WW1), minVal(WW2), minVal(WW3), minVal(VV) (lMag) */
    double umach;     /* UMACH (3dmhdsub.h:42-44)                                 */
    double dt;        /* SFF * min(...) (3dmhdsub.h:101-105)                      */
} mhd3d_cfl_info;

/* Kernel timing gathered with HIP events on the context stream. */
typedef struct mhd3d_timing {
    double stage_ms_total;   /* summed duration of the stage kernels           */
    int64_t stage_launches;  /* number of stage-kernel launches timed          */
    double step_ms_total;    /* summed duration of whole steps                 */
    int64_t steps;
} mhd3d_timing;

/* Create a context: validates the descriptor, copies the metric arrays and
 * allocates the device registers (U ping-pong + Z).  Replaces the static
 * storage of 3dmhd.cpp:24-79 and the setup half of main(). */
int mhd3d_create(const mhd3d_desc* desc, mhd3d_ctx** out);
int mhd3d_destroy(mhd3d_ctx* ctx);

/* Upload / download the state.  `fields[f]` points at a host array
 * [nPz+2][nPy+2][nPx+2]; the magnetic slots are read only with lMag.
 * Upload takes the interior (x,y) of every plane including the two z ghost
 * planes; download writes the interior and refills the x/y ghosts periodically
 * as communicate() does (3dmhdmpi.h:59-142).  Replaces direct access to the
 * globals of 3dmhddata.h:12-17. */
int mhd3d_set_state(mhd3d_ctx* ctx, const double* const* fields);
int mhd3d_get_state(mhd3d_ctx* ctx, double* const* fields);

/* Advance `nsteps` RK3 steps: step() of 3dmhdsub.h:31-153, nsteps times.
 * fixed_dt == 0: each step computes DT from the CFL condition (TSTEP) and, on
 * return, *dt_inout holds the DT of the last step.  fixed_dt != 0: every step
 * uses *dt_inout.  Returns MHD3D_ENONFINITE when the state is not finite
 * after the steps, MHD3D_EHALTED when the DT guard stopped the run (the state
 * is that after the step whose DT fell below the threshold; mhd3d_clock
 * gives the steps done). */
int mhd3d_step(mhd3d_ctx* ctx, int nsteps, int fixed_dt, double* dt_inout);

/* mhd3d_step without device allocation.  No counterpart in the reference
 * (its step(), 3dmhdsub.h:31, works on static arrays): for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Same arguments, results
 * and status codes as mhd3d_step, bit for bit; the context and its state
 * are the caller's (created beforehand).  Its TSTEP partial records live in
 * one process-wide array of 32768 records (1.5 MiB) per GPU instead of the
 * context's: grids that need more (nPx*nPy beyond about 1M points) get
 * MHD3D_EINVAL with a message, as does a call with kernel timing enabled or
 * from a device other than the context's.  Calls on one GPU share that
 * array: calls that overlap in time (other streams or threads) race. */
int mhd3d_step_noalloc(mhd3d_ctx* ctx, int nsteps, int fixed_dt, double* dt_inout);

/* Autotuner of a single domain (no counterpart in the reference): times
 * `steps` CFL steps of each streaming-kernel variant -- register budget
 * (1 or 2 waves per SIMD) x schedule (4) x launch size (one round of the
 * resident blocks, or z chunks of at most 64 or 32 planes): 24 candidates --
 * on the current state, restores the state, the clock and the TSTEP after
 * each trial and keeps the fastest.  Results do not depend on the variant.
 * report: NULL or room for 4*24 doubles (waves, sched, planes per chunk
 * (0 = one round), ms per step; -1 for a trial the DT guard stopped);
 * *ncand = trials.  MHD3D_EUNSUPPORTED for the point kernel and for the
 * off-default branches (their state between steps, one variant). */
int mhd3d_autotune(mhd3d_ctx* ctx, int steps, double* report, int* ncand);
/* The variant in use: waves per SIMD, schedule bits, blocks per launch. */
int mhd3d_get_variant(mhd3d_ctx* ctx, int* waves, int* sched, int* blocks);

/* DT guard of the reference's driver (3dmhd.cpp:292-330): after any step
 * with DT < dt_min the run halts -- later steps do nothing until the next
 * set_state.  dt_min = 0 (the default) disables it; MHD3D_REFERENCE_DT_MIN is
 * the reference's threshold.  Checked on the device, without host syncs. */
int mhd3d_set_dt_guard(mhd3d_ctx* ctx, double dt_min);

/* fluxes() of 3dmhdsub.h:194-696 on the current state, written to host
 * arrays in the field layout: FU,FV,FW,FR,FT and, with lMag, WW1,WW2,WW3
 * (interior incl. walls; ghosts zero).  Pure: the state is not modified
 * (the reference inverts RO in place and uses RU..TT as scratch). */
int mhd3d_fluxes(mhd3d_ctx* ctx, double* const* flux);

/* One RK3 stage in isolation, for parity tests of the fused update:
 * stage 1..3 of 3dmhdsub.h:116-146 (ZETA/fluxes/GAM/BCON/communicate). */
int mhd3d_stage(mhd3d_ctx* ctx, int stage, double dt);

/* BCON() (3dmhdsub.h:720-873) applied to the current state. */
int mhd3d_bcon(mhd3d_ctx* ctx);

/* TSTEP block of step() (3dmhdsub.h:36-106) on the current state. */
int mhd3d_cfl(mhd3d_ctx* ctx, mhd3d_cfl_info* out);

/* Global diagnostics in FP64: total mass sum(RO*dV), total energy
 * sum((0.5*|RU,RV,RW|^2/RO + CV*RO*TT [+ OBETA*|B|^2/2])*dV) and max|div B|
 * over the interior including walls, with the cell volume
 * dV = DDX[x]*DDY[y]*DDZ[z] of the metric (3dmhddata.h:20-22; DDX = 1 when
 * nPx == 1). */
int mhd3d_diagnostics(mhd3d_ctx* ctx, double out[3]);

/* State of the RLAX relaxation, the function statics of fluxes()
 * (3dmhdsub.h:195-198): out[0] = TFAC, out[1] = TTM_MAX, out[2] = TTM_MIN,
 * *icall = ICALL (fluxes() calls so far), *exited = 1 after the reference's
 * exit point (then every step call returns MHD3D_EEXIT and the state is that
 * at the start of the step in which the exit fell, or, for an exit in a
 * later stage of the FP64 form, that stage's input), else 0.  A context
 * starts as a fresh process does: TFAC = 1, TTM_MAX = -1e99, TTM_MIN = 1e99,
 * ICALL = 0; every fluxes() evaluation (each RK stage, mhd3d_fluxes)
 * advances it as the reference's call does.  mhd3d_set_state and
 * mhd3d_init_stratified clear the exit (the domain runs again); the other
 * statics persist, as a process's function statics do. */
int mhd3d_relax_state(mhd3d_ctx* ctx, double out[3], int64_t* icall, int* exited);

/* Simulation clock: nIt, TIMT, TIMC (3dmhdsub.h:150-152). */
int mhd3d_clock(mhd3d_ctx* ctx, int64_t* nIt, double* timt, double* timc);

/* Event timing of the stage kernels (enable before the timed region). */
int mhd3d_timing_enable(mhd3d_ctx* ctx, int on);
int mhd3d_timing_read(mhd3d_ctx* ctx, mhd3d_timing* out);

int mhd3d_synchronize(mhd3d_ctx* ctx);
/* Text of the last error on this context (or of the last failed create). */
const char* mhd3d_last_error(const mhd3d_ctx* ctx);
/* Bytes of device memory the context holds. */
int64_t mhd3d_device_bytes(const mhd3d_ctx* ctx);

/* Device-side synthetic stratified state (no counterpart in the reference,
 * whose initial state comes from Static(), 3dmhdphsic.h:549-763): RO, TT from
 * the z profiles RO_z[nz], TT_z[nz] (nz = nPz + 2, ghost planes included);
 * RU, RV, RW = RO * amp * u on the planes between the walls; with lMag
 * BX = b0 + b_amp*u, BY, BZ = b_amp*u on planes 1..nPz, BZ = 0 on the walls,
 * ghost planes copies of the walls; u in [-1, 1) a counter-based hash of
 * (seed, field, z, y, x).  For callers that must not ship fields from the
 * host (the request loop of a server). */
int mhd3d_init_stratified(mhd3d_ctx* ctx, const double* TT_z, const double* RO_z, uint64_t seed,
                          double amp, double b0, double b_amp);

/* ------------------------------------------------------------------------
 * Ensembles: n independent domains (sizes, flags and parameters may differ;
 * one precision form for all) advanced together -- one persistent launch per
 * RK3 stage for the planes between the walls and one for the wall planes of
 * every member, each member with its own CFL time step.  A member advances
 * bit-for-bit as the same domain does alone through mhd3d_step.  Members of
 * either form may take the off-default branches (iD, lRem, lShr, iTCon 2,
 * the TP plume, RLAX), except the heat loss and the plume with TC != 0,
 * whose factors of TIMT the host forms before every step
 * (MHD3D_EUNSUPPORTED); a member that reaches an RLAX exit point stops there
 * and mhd3d_ensemble_step returns MHD3D_EEXIT.  No counterpart in the reference (its step() advances
 * one domain, 3dmhdsub.h:31).
 * ------------------------------------------------------------------------ */
typedef struct mhd3d_ensemble mhd3d_ensemble;

int mhd3d_ensemble_create(const mhd3d_desc* const* descs, int n, mhd3d_ensemble** out);
int mhd3d_ensemble_destroy(mhd3d_ensemble* e);
const char* mhd3d_ensemble_last_error(const mhd3d_ensemble* e);   /* e == NULL: last create error */
int mhd3d_ensemble_size(const mhd3d_ensemble* e);
int64_t mhd3d_ensemble_cells(const mhd3d_ensemble* e);
/* fields as for mhd3d_set_state / mhd3d_get_state, for member `member`;
 * set_state also resets that member's clock (its DT guard stays) */
int mhd3d_ensemble_set_state(mhd3d_ensemble* e, int member, const double* const* fields);
int mhd3d_ensemble_get_state(mhd3d_ensemble* e, int member, double* const* fields);
/* nsteps RK3 steps of every member, DT from each member's CFL (TSTEP) */
int mhd3d_ensemble_step(mhd3d_ensemble* e, int nsteps);
int mhd3d_ensemble_clock(mhd3d_ensemble* e, int member, int64_t* nIt, double* timt,
                         double* timc, double* dt_last);
/* on != 0: each step is replayed from a captured hipGraph */
int mhd3d_ensemble_graph(mhd3d_ensemble* e, int on);
/* kernel variant: waves (1, 2) = register budget, sched bits (1: one barrier
 * per plane, 2: deferred RK update), chunk = planes per work item (0 = auto);
 * the arithmetic does not depend on it */
int mhd3d_ensemble_set_variant(mhd3d_ensemble* e, int waves, int sched, int chunk);
int mhd3d_ensemble_get_variant(const mhd3d_ensemble* e, int* waves, int* sched, int* chunk);
int mhd3d_ensemble_synchronize(mhd3d_ensemble* e);
/* the DT guard (mhd3d_set_dt_guard) for every member, or for one member; a
 * halted member stays as it is while the others advance */
int mhd3d_ensemble_set_dt_guard(mhd3d_ensemble* e, double dt_min);
int mhd3d_ensemble_set_dt_guard_member(mhd3d_ensemble* e, int member, double dt_min);
int mhd3d_ensemble_halted(mhd3d_ensemble* e, int member, int* halted);
/* Autotuner: times `steps` steps of every variant (waves x sched x chunk, 32
 * candidates) on the current states, restores states and clocks after each
 * trial and keeps the fastest.  report: NULL or room for 4*32 doubles
 * (waves, sched, chunk, ms per step); *ncand = number of trials. */
int mhd3d_ensemble_autotune(mhd3d_ensemble* e, int steps, double* report, int* ncand);

/* ------------------------------------------------------------------------
 * Request server (no counterpart in the reference): runs arrive at any time
 * and are admitted first-come-first-served while their device memory fits
 * the HBM budget; their state is built on the device (as
 * mhd3d_init_stratified) and kept there; each tick advances every admitted
 * run by `quantum_steps` steps in one ensemble, each run stopping on the
 * device at its own step count (continuous batching: runs join and leave
 * between ticks; runs that arrive while a quantum runs are set up on the
 * host and device meanwhile and join at the next tick).  A run's result is
 * the same as the same run advanced alone.  A run whose state turns
 * non-finite is evicted (MHD3D_REQ_NONFINITE).  Times are seconds since the
 * server's creation.
 * ------------------------------------------------------------------------ */
typedef struct mhd3d_server mhd3d_server;

enum {
    MHD3D_REQ_EXITED = -3,   /* stopped where the reference's program exits
                                (the RLAX relaxation, MHD3D_EEXIT)          */
    MHD3D_REQ_NONFINITE = -2,/* evicted: its state held NaN/Inf after a tick
                                (the reference's driver would keep stepping
                                it, its minima ignore NaN, 3dmhdutils.h:61-79) */
    MHD3D_REQ_FAILED = -1,   /* halted by the reference's DT guard (every
                                run is admitted with MHD3D_REFERENCE_DT_MIN) */
    MHD3D_REQ_QUEUED = 0,
    MHD3D_REQ_RUNNING = 1,
    MHD3D_REQ_DONE = 2,      /* state on the device until mhd3d_server_result */
    MHD3D_REQ_RELEASED = 3
};

typedef struct mhd3d_request_info {
    int32_t status;
    int32_t steps_done;
    double t_submit, t_admit, t_finish;   /* -1 while not reached */
    double dt_last;
} mhd3d_request_info;

int mhd3d_server_create(int precision, int64_t hbm_budget_bytes, int quantum_steps,
                        mhd3d_server** out);
int mhd3d_server_destroy(mhd3d_server* s);
const char* mhd3d_server_last_error(const mhd3d_server* s);
double mhd3d_server_now(const mhd3d_server* s);
int64_t mhd3d_server_bytes_in_use(const mhd3d_server* s);
int64_t mhd3d_server_cell_steps(const mhd3d_server* s);
/* desc and its arrays are copied; the initial state is that of
 * mhd3d_init_stratified(TT_z, RO_z, seed, amp, b0, b_amp); the branches an
 * ensemble member may take are accepted (MHD3D_EUNSUPPORTED otherwise) */
int mhd3d_server_submit(mhd3d_server* s, const mhd3d_desc* desc, const double* TT_z,
                        const double* RO_z, uint64_t seed, double amp, double b0, double b_amp,
                        int nsteps, int64_t* id);
int mhd3d_server_tick(mhd3d_server* s, int* n_active, int* n_queued);
int mhd3d_server_poll(mhd3d_server* s, int64_t id, mhd3d_request_info* info);
/* copies the final state of a finished run (fields may be NULL; only DONE
 * runs copy) and releases the device memory of any finished run */
int mhd3d_server_result(mhd3d_server* s, int64_t id, double* const* fields);

/* nPx, nPy, nPz and the number of fields of a context. */
int mhd3d_geometry(const mhd3d_ctx* ctx, int32_t out[4]);

/* The reference's dump/restart files (3dmhdio.h:20-72).  .dat0: dump idx =
 * records n_f*idx .. n_f*idx+n_f-1 of nPx*nPy*nPz doubles, fields RU, RV, RW,
 * TT, RO [, BX, BY, BZ], planes 1..nPz x rows 1..nPy x columns 1..nPx, x
 * fastest.  mhd3d_load_data keeps the context's z ghost planes (the
 * reference's restart, 3dmhd.cpp:112-133, leaves them as they were) and
 * fills x/y ghosts periodically.  .par: record idx = 96 doubles (par1:
 * ipar[0..31] then par[0..63], with nIt at [14], DT at [62], TIMT/TIMI at
 * [63], 0/TIMC at [64]; 3dmhd.cpp:271-289).  Other records of a file are
 * kept; missing files are created. */
int mhd3d_dump_data(mhd3d_ctx* ctx, const char* path, int idx);
int mhd3d_load_data(mhd3d_ctx* ctx, const char* path, int idx);
int mhd3d_dump_param(const char* path, int idx, const double* par1);
int mhd3d_load_param(const char* path, int idx, double* par1);

/* The reference's initial state Static() (3dmhdphsic.h:549-763) of a run on
 * nPEy x nPEz ranks, assembled into one domain of nPEz*(nPz_global/nPEz)
 * planes: the polytrope of derivs() integrated by bsstep on the global z grid
 * (nGrid = 0, PZP = 0), each plane as its rank takes it (the last z rank
 * aligns with the top of the profile), TT times 1 + AMPT*(RAN2 - 1/2) drawn
 * in rank 0's loop order (IDUM = -62659), RU = RV = RW = 0; x/y ghosts
 * periodic.  par = the reference's par[64] (THETA = par[2], GRAV = par[3],
 * zMax = par[18], AMPT = par[19]).  fields: RU, RV, RW, RO, TT (and BX, BY,
 * BZ with lMag) host arrays of [nPEz*nRz+2][nPy+2][nPx+2].  lMag: the
 * AMPB Gaussian BX layer on each rank's planes 2..nRz+1 with RO lowered by
 * BX^2/TT*OBETA (AMPB = par[20], BFH = par[21], BZP = par[22], OBETA =
 * 2.0f/par[7]; 3dmhdphsic.h:682-699), or B = 0 (tube() with nTube = 0).
 * Host-only; needs no GPU. */
int mhd3d_static_state(int nPx, int nPy, int nPz_global, int nPEy, int nPEz, const double* par,
                       int mag, double* const* fields);

/* Static() with the lShr and lRem switches as well (3dmhdphsic.h:565-616,
 * 737-742): lShr starts from T = RO = 1 and puts the RAN2 perturbation into
 * RV (RV + (1 + AMPT*(RAN2 - 1/2)) - 1); lRem takes M. Rempel's background
 * (getbackground/ODEINT/RKQC/RK4/derivs1/kappa, :357-522, REPR = par[0]*par[1])
 * and also returns the conductivity profile it sets, RKAPA and DKAPA
 * ([nPEz*nRz+2], the rank assembly of RK*8.07f*THETA*REPR and its one-sided /
 * central differences), and walls = {TU, DZTU, DZTB} for BCON; RKAPA, DKAPA
 * and walls may be NULL without lRem.  lShr with lRem: MHD3D_EUNSUPPORTED. */
int mhd3d_static_state_flags(int nPx, int nPy, int nPz_global, int nPEy, int nPEz,
                             const double* par, int mag, int shr, int rem,
                             double* const* fields, double* RKAPA, double* DKAPA, double* walls);

/* Measurement support (no counterpart in the reference): `reps` launches of
 * a plain device copy of `bytes` bytes between two fresh buffers, to
 * calibrate a profiler's HBM byte counters against a known traffic;
 * *ms_per_copy gets the event-timed average. */
int mhd3d_calibration_copy(int64_t bytes, int reps, double* ms_per_copy);

/* Measurement support (no counterpart in the reference): the streaming
 * ceiling of the memory system for the read/write mixes of the RK stages --
 * `reps` launches of a tuned kernel (16-byte vector loads, 4 in flight per
 * thread and array) reading nr (1-2) arrays and writing nw (0-2) arrays of
 * `bytes_per_array` bytes each; *ms_per_launch gets the event-timed average,
 * so the rate is (nr + nw) * bytes_per_array / time. */
int mhd3d_stream_ceiling(int64_t bytes_per_array, int nr, int nw, int reps, double* ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif /* MHD3D_H_ */
