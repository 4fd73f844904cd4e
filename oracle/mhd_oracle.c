/*
 * mhd_oracle.c — CPU restatement of the reference RK3 timestep.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path; see mhd_oracle.h).
 *
 * Every operator below is one whole-array sweep, in the order of the
 * reference's fluxes() (3dmhdsub.h:194-696), with the same left-to-right
 * product order as the operator templates of 3dmhdutils.h:168-239:
 *   d1        ((f[+1]-f[-1]) * du) * h * s                       (:195-199)
 *   d1_mul    (((f[+1]-f[-1]) * du) * h * g) * s                 (:201-205)
 *   d1_mul_az ((((f[+1]-f[-1]) * du) * h * g) * a[z]) * s         (:207-211)
 *   d2        (((df*d2u)*h) + ((f[+1]-2f)+f[-1])*du^2*h2) * s     (:215-220)
 *   d2_mul_az ((d2expr * g) * a[z]) * s                           (:229-234)
 *   mul_d1    ((F[+1]G[+1] - F[-1]G[-1]) * du) * h                 (:236-239)
 *   mul2/mul3/base: (a*b)*s, ((a*b)*c)*s, a*s                     (:169-187)
 * SET/ADD/SUB accumulate as oper<> (3dmhdutils.h:126-129).
 * Compile with -ffp-contract=off: no fused multiply-adds, like the x86-64
 * baseline build of the reference.
 *
 * Differences from the reference that do not change the state after a step:
 *  - RO is not inverted in place (3dmhdsub.h:268); 1/RO lives in RI.  Only
 *    wall-plane fluxes that BCON overwrites read the z-ghost planes, so the
 *    reference's alternating rho <-> 1/rho z-ghosts do not reach the state.
 *  - The lMag block's scratch arrays are S1..S4 instead of RU,RV,RW,TT
 *    (3dmhdsub.h:548-694), which step_block_GAM refills anyway.
 *  - Both walls belong to the one domain: the one-sided stencils of
 *    3dmhdsub.h:572-584,640-654,673-687 apply at both walls (the `else if`
 *    of the reference skips the bottom one when nPEz == 1).
 */
#include "mhd_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Planted-mistake switches for the tests that show the mathematical checks
 * catch a dropped term, a wrong sign, a transposed operand or a wrong index
 * (tests/test_oracle_mutants.py builds the library with -DORC_MUTANT=k).
 * The default build has ORC_MUTANT == 0 and none of them is active. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define MUT(k) (ORC_MUTANT == (k))

enum { SET = 0, ADD = 1, SUB = 2 };
enum { DZ = 0, DY = 1, DX = 2 };

typedef struct { int zl, zr, yl, yr, xl, xr; } rng_t;

struct orc_ctx {
    mhd3d_desc d;
    int nx, ny, nz;
    size_t n;
    /* BIG (3dmhddata.h:12-19) */
    double *RU, *RV, *RW, *RO, *TT, *UU, *VV, *WW, *FU, *FV, *FW, *FR, *FT;
    double *ZRU, *ZRV, *ZRW, *ZRO, *ZTT, *WW1, *WW2, *WW3, *BX, *BY, *BZ;
    double *ZBX, *ZBY, *ZBZ;
    double *RI;                     /* 1/RO over the extended range */
    double *S1, *S2, *S3, *S4;      /* lMag scratch                 */
    /* AJACOBI + CPEN */
    double *dxxdx, *d2xxdx2, *DDX, *dyydy, *d2yydy2, *DDY;
    double *dzzdz, *d2zzdz2, *DDZ, *RKAPA, *DKAPA;
    double *EXX, *WYY, *ZEE;        /* coordinates (off-default branches)   */
    /* function-local statics of fluxes()/horizontal_mean (3dmhdsub.h:200,699) */
    double *ROM, *TTM, *HRAD, *WWZ;
    /* RLAX relaxation (3dmhdsub.h:195-200): the statics ICALL, TTM_MIN,
     * TTM_MAX, TFAC, the profiles FCONM, FRADM, CORRECT over the whole domain,
     * and the exit point the reference's program stops at (1: "yk not
     * implement", :427-439; 2: "Relax: check boundary temperature", :344,416) */
    double *FCONM, *FRADM, *CORRECT;
    double TFAC, TTM_MAX, TTM_MIN;
    long long ICALL;
    int relax_exit;
    /* CTIM / ITER */
    double DT, TIMT, TIMC, TIMI, UMACH;
    long long nIt;
};

/* ------------------------------------------------------------------------ */
/* sweep machinery                                                          */

#define NX_ c->nx
#define NY_ c->ny

#define LOOP(R, STMT)                                                              \
    do {                                                                           \
        const int zl_ = (R).zl, zr_ = (R).zr, yl_ = (R).yl, yr_ = (R).yr;          \
        const int xl_ = (R).xl, xr_ = (R).xr, nx_ = NX_, ny_ = NY_;                \
        _Pragma("omp parallel for collapse(2) schedule(static)")                   \
        for (int z = zl_; z < zr_; ++z)                                            \
            for (int y = yl_; y < yr_; ++y) {                                      \
                const size_t r0_ = ((size_t)z * ny_ + y) * nx_;                    \
                for (int x = xl_; x < xr_; ++x) {                                  \
                    const size_t i = r0_ + x;                                      \
                    (void)i;                                                       \
                    STMT;                                                          \
                }                                                                  \
            }                                                                      \
    } while (0)

#define LOOP_OP(op, R, LHS, RHS)                      \
    do {                                              \
        if ((op) == SET) LOOP(R, LHS = (RHS));        \
        else if ((op) == ADD) LOOP(R, LHS += (RHS));  \
        else LOOP(R, LHS -= (RHS));                   \
    } while (0)

/* per-dimension stride, grid factor and metric arrays (3dmhdutils.h:136-166) */
#define DIMSETUP(dim)                                                                   \
    const ptrdiff_t s = (dim) == DX ? 1 : (dim) == DY ? (ptrdiff_t)c->nx                \
                                                      : (ptrdiff_t)c->nx * c->ny;       \
    const double h = (dim) == DX ? c->d.hx : (dim) == DY ? c->d.hy : c->d.hz;           \
    const double h2 = (dim) == DX ? c->d.h2x : (dim) == DY ? c->d.h2y : c->d.h2z;       \
    const double* du = (dim) == DX ? c->dxxdx : (dim) == DY ? c->dyydy : c->dzzdz;      \
    const double* d2u = (dim) == DX ? c->d2xxdx2 : (dim) == DY ? c->d2yydy2 : c->d2zzdz2; \
    (void)h2; (void)d2u;
#define MI(dim) ((dim) == DX ? x : (dim) == DY ? y : z)

/* metric index of a first difference along z (mutant 19 shifts it by one plane) */
#define DU1(dim) (MI(dim) + ((MUT(19) && (dim) == DZ) ? 1 : 0))

static void f_d1(orc_ctx* c, int op, int dim, rng_t r, double* t, const double* f, double sc)
{
    DIMSETUP(dim);
    LOOP_OP(op, r, t[i], ((f[i + s] - f[i - s]) * du[DU1(dim)]) * h * sc);
}

static void f_d1_mul(orc_ctx* c, int op, int dim, rng_t r, double* t, const double* f,
                     const double* g, double sc)
{
    DIMSETUP(dim);
    LOOP_OP(op, r, t[i], ((f[i + s] - f[i - s]) * du[MI(dim)]) * h * g[i] * sc);
}

static void f_d1_mul_az(orc_ctx* c, int op, int dim, rng_t r, double* t, const double* f,
                        const double* g, const double* az, double sc)
{
    DIMSETUP(dim);
    LOOP_OP(op, r, t[i], ((f[i + s] - f[i - s]) * du[MI(dim)]) * h * g[i] * az[z] * sc);
}

/* mutants 16 (metric term of D2z dropped) and 20 (dzzdz not squared) */
#define D2EXPR(f)                                                              \
    ((f[i + s] - f[i - s]) * ((MUT(16) && (dim) == DZ) ? 0.0 : d2u[MI(dim)]) * h + \
     ((f[i + s] - 2 * f[i]) + f[i - s]) *                                      \
         ((MUT(20) && (dim) == DZ) ? du[MI(dim)] : (du[MI(dim)] * du[MI(dim)])) * h2)

static void f_d2(orc_ctx* c, int op, int dim, rng_t r, double* t, const double* f, double sc)
{
    DIMSETUP(dim);
    LOOP_OP(op, r, t[i], D2EXPR(f) * sc);
}

static void f_d2_mul_az(orc_ctx* c, int op, int dim, rng_t r, double* t, const double* f,
                        const double* g, const double* az, double sc)
{
    DIMSETUP(dim);
    LOOP_OP(op, r, t[i], D2EXPR(f) * g[i] * az[z] * sc);
}

static void f_d2_mul(orc_ctx* c, int op, int dim, rng_t r, double* t, const double* f,
                     const double* g, double sc)
{
    DIMSETUP(dim);
    LOOP_OP(op, r, t[i], D2EXPR(f) * g[i] * sc);
}

/* func_mul_extend / func_base_extend along z (3dmhdutils.h:173-187) */
static void f_mul_ext(orc_ctx* c, int op, rng_t r, double* t, const double* a, const double* az,
                      double sc)
{
    LOOP_OP(op, r, t[i], a[i] * az[z] * sc);
}

static void f_base_ext(orc_ctx* c, int op, rng_t r, double* t, const double* az, double sc)
{
    LOOP_OP(op, r, t[i], az[z] * sc);
}

static void f_mul_d1(orc_ctx* c, int op, int dim, rng_t r, double* t, const double* f,
                     const double* g)
{
    DIMSETUP(dim);
    LOOP_OP(op, r, t[i], (f[i + s] * g[i + s] - f[i - s] * g[i - s]) * du[MI(dim)] * h);
}

static void f_base(orc_ctx* c, int op, rng_t r, double* t, const double* a, double sc)
{
    LOOP_OP(op, r, t[i], a[i] * sc);
}

static void f_mul2(orc_ctx* c, int op, rng_t r, double* t, const double* a, const double* b,
                   double sc)
{
    LOOP_OP(op, r, t[i], a[i] * b[i] * sc);
}

static void f_mul3(orc_ctx* c, int op, rng_t r, double* t, const double* a, const double* b,
                   const double* e, double sc)
{
    LOOP_OP(op, r, t[i], a[i] * b[i] * e[i] * sc);
}

/* scale_add / copy of 3dmhdutils.h:53-59 */
static void f_scale_add(orc_ctx* c, rng_t r, double* t, const double* a, const double* b,
                        double sc)
{
    LOOP(r, t[i] = a[i] + sc * b[i]);
}

static void f_copy(orc_ctx* c, rng_t r, double* t, const double* a)
{
    LOOP(r, t[i] = a[i]);
}

/* ------------------------------------------------------------------------ */

#define IX(z, y, x) (((size_t)(z) * c->ny + (y)) * c->nx + (x))

static rng_t R_in(const orc_ctx* c) { rng_t r = {1, c->nz - 1, 1, c->ny - 1, 1, c->nx - 1}; return r; }
static rng_t R_ext(const orc_ctx* c) { rng_t r = {0, c->nz, 0, c->ny, 0, c->nx}; return r; }
static rng_t R_exy(const orc_ctx* c) { rng_t r = {1, c->nz - 1, 0, c->ny, 0, c->nx}; return r; }
static rng_t R_ex(const orc_ctx* c) { rng_t r = {1, c->nz - 1, 1, c->ny - 1, 0, c->nx}; return r; }
static rng_t R_ey(const orc_ctx* c) { rng_t r = {1, c->nz - 1, 0, c->ny, 1, c->nx - 1}; return r; }

static double* dup_arr(const double* src, int n)
{
    double* p = (double*)calloc((size_t)n, sizeof(double));
    if (p && src) memcpy(p, src, (size_t)n * sizeof(double));
    return p;
}

orc_ctx* orc_create(const mhd3d_desc* d)
{
    if (!d || d->nPx < 1 || d->nPy < 1 || d->nPz < 4) return NULL;
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    if (!c) return NULL;
    c->d = *d;
    c->nx = d->nPx + 2;
    c->ny = d->nPy + 2;
    c->nz = d->nPz + 2;
    c->n = (size_t)c->nx * c->ny * c->nz;
    double** big[] = {&c->RU, &c->RV, &c->RW, &c->RO, &c->TT, &c->UU, &c->VV, &c->WW,
                      &c->FU, &c->FV, &c->FW, &c->FR, &c->FT, &c->ZRU, &c->ZRV, &c->ZRW,
                      &c->ZRO, &c->ZTT, &c->WW1, &c->WW2, &c->WW3, &c->BX, &c->BY, &c->BZ,
                      &c->ZBX, &c->ZBY, &c->ZBZ, &c->RI, &c->S1, &c->S2, &c->S3, &c->S4};
    for (size_t k = 0; k < sizeof(big) / sizeof(big[0]); ++k) {
        *big[k] = (double*)calloc(c->n, sizeof(double));
        if (!*big[k]) { orc_destroy(c); return NULL; }
    }
    c->dxxdx = dup_arr(d->dxxdx, c->nx);
    c->d2xxdx2 = dup_arr(d->d2xxdx2, c->nx);
    c->DDX = dup_arr(d->DDX, c->nx);
    c->dyydy = dup_arr(d->dyydy, c->ny);
    c->d2yydy2 = dup_arr(d->d2yydy2, c->ny);
    c->DDY = dup_arr(d->DDY, c->ny);
    c->dzzdz = dup_arr(d->dzzdz, c->nz);
    c->d2zzdz2 = dup_arr(d->d2zzdz2, c->nz);
    c->DDZ = dup_arr(d->DDZ, c->nz);
    c->RKAPA = dup_arr(d->RKAPA, c->nz);
    c->DKAPA = dup_arr(d->DKAPA, c->nz);
    c->EXX = dup_arr(d->EXX, c->nx);
    c->WYY = dup_arr(d->WYY, c->ny);
    c->ZEE = dup_arr(d->ZEE, c->nz);
    c->ROM = dup_arr(NULL, c->nz);
    c->TTM = dup_arr(NULL, c->nz);
    c->HRAD = dup_arr(NULL, c->nz);
    c->WWZ = dup_arr(NULL, c->nz);
    c->FCONM = dup_arr(NULL, c->nz);
    c->FRADM = dup_arr(NULL, c->nz);
    c->CORRECT = dup_arr(NULL, c->nz);
    c->TFAC = 1.0;                              /* 3dmhdsub.h:196-198 */
    c->TTM_MAX = -1e99;
    c->TTM_MIN = 1e99;
    c->d.EXX = c->d.WYY = c->d.ZEE = NULL;   /* the copies above are used */
    c->TIMI = d->TIMI;
    c->TIMT = d->TIMI;
    c->TIMC = 0.0;
    return c;
}

void orc_destroy(orc_ctx* c)
{
    if (!c) return;
    double* all[] = {c->RU, c->RV, c->RW, c->RO, c->TT, c->UU, c->VV, c->WW, c->FU, c->FV,
                     c->FW, c->FR, c->FT, c->ZRU, c->ZRV, c->ZRW, c->ZRO, c->ZTT, c->WW1,
                     c->WW2, c->WW3, c->BX, c->BY, c->BZ, c->ZBX, c->ZBY, c->ZBZ, c->RI,
                     c->S1, c->S2, c->S3, c->S4, c->dxxdx, c->d2xxdx2, c->DDX, c->dyydy,
                     c->d2yydy2, c->DDY, c->dzzdz, c->d2zzdz2, c->DDZ, c->RKAPA, c->DKAPA,
                     c->EXX, c->WYY, c->ZEE, c->ROM, c->TTM, c->HRAD, c->WWZ,
                     c->FCONM, c->FRADM, c->CORRECT};
    for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) free(all[k]);
    free(c);
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

static double* lookup(orc_ctx* c, const char* nm)
{
    struct { const char* k; double* v; } t[] = {
        {"RU", c->RU}, {"RV", c->RV}, {"RW", c->RW}, {"RO", c->RO}, {"TT", c->TT},
        {"BX", c->BX}, {"BY", c->BY}, {"BZ", c->BZ}, {"UU", c->UU}, {"VV", c->VV},
        {"WW", c->WW}, {"FU", c->FU}, {"FV", c->FV}, {"FW", c->FW}, {"FR", c->FR},
        {"FT", c->FT}, {"WW1", c->WW1}, {"WW2", c->WW2}, {"WW3", c->WW3},
        {"ZRU", c->ZRU}, {"ZRV", c->ZRV}, {"ZRW", c->ZRW}, {"ZRO", c->ZRO},
        {"ZTT", c->ZTT}, {"ZBX", c->ZBX}, {"ZBY", c->ZBY}, {"ZBZ", c->ZBZ}, {"RI", c->RI}};
    for (size_t k = 0; k < sizeof(t) / sizeof(t[0]); ++k)
        if (strcmp(t[k].k, nm) == 0) return t[k].v;
    return NULL;
}

int orc_set(orc_ctx* c, const char* name, const double* a)
{
    double* p = lookup(c, name);
    if (!p || !a) return -1;
    memcpy(p, a, c->n * sizeof(double));
    return 0;
}

int orc_get(orc_ctx* c, const char* name, double* a)
{
    double* p = lookup(c, name);
    if (!p || !a) return -1;
    memcpy(a, p, c->n * sizeof(double));
    return 0;
}

void orc_set_time(orc_ctx* c, double dt, double timt)
{
    c->DT = dt;
    c->TIMT = timt;
    c->TIMC = timt - c->TIMI;
}

int orc_relax_state(orc_ctx* c, double out[3], long long* icall, int* exited, double* correct,
                    double* fconm, double* fradm)
{
    if (!c) return -1;
    if (out) {
        out[0] = c->TFAC;
        out[1] = c->TTM_MAX;
        out[2] = c->TTM_MIN;
    }
    if (icall) *icall = c->ICALL;
    if (exited) *exited = c->relax_exit;
    const size_t b = (size_t)c->nz * sizeof(double);
    if (correct) memcpy(correct, c->CORRECT, b);
    if (fconm) memcpy(fconm, c->FCONM, b);
    if (fradm) memcpy(fradm, c->FRADM, b);
    return 0;
}

void orc_clock(orc_ctx* c, long long* nIt, double* timt, double* timc)
{
    if (nIt) *nIt = c->nIt;
    if (timt) *timt = c->TIMT;
    if (timc) *timc = c->TIMC;
}

/* ------------------------------------------------------------------------ */
/* communicate(): x and y periodic, z ghosts untouched (3dmhdmpi.h:32-143    */
/* with nPEy = 1 and one z domain).                                          */

static void periodic_fill(orc_ctx* c, double* a)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z) {
        for (int y = 1; y < ny - 1; ++y) {              /* :118-121 */
            a[IX(z, y, 0)] = a[IX(z, y, nx - 2)];
            a[IX(z, y, nx - 1)] = a[IX(z, y, 1)];
        }
        for (int x = 0; x < nx; ++x) {                  /* :65-68, :135-142 */
            a[IX(z, 0, x)] = a[IX(z, ny - 2, x)];
            a[IX(z, ny - 1, x)] = a[IX(z, 1, x)];
        }
    }
}

void orc_communicate(orc_ctx* c)
{
    periodic_fill(c, c->RU);
    periodic_fill(c, c->RV);
    periodic_fill(c, c->RW);
    periodic_fill(c, c->TT);
    periodic_fill(c, c->RO);
    if (c->d.flag_lMag) {
        periodic_fill(c, c->BX);
        periodic_fill(c, c->BY);
        periodic_fill(c, c->BZ);
    }
}

/* ------------------------------------------------------------------------ */
/* horizontal_mean — 3dmhdsub.h:698-718 (nGrid 0, one y rank): the running    */
/* sum over the plane interior, y outer, x inner, divided by float(nPx*nPy). */

static void hmean(orc_ctx* c, double* varM, const double* var)
{
    const int nx = c->nx, ny = c->ny;
    const double den = (float)(c->d.nPx * c->d.nPy);
    for (int z = 0; z < c->nz; ++z) {
        double w = 0;
        for (int y = 1; y < ny - 1; y++)
            for (int x = 1; x < nx - 1; x++) w += var[IX(z, y, x)];
        varM[z] = w / den;
    }
}

static double sq(double v) { return v * v; }   /* square() of 3dmhdutils.h:213 */

/* -4.0f*log(2.0f): the reference's translation unit includes <math.h>
 * (3dmhdphsic.h:11), whose C++ wrapper puts the float overload of log() in
 * the global namespace, so log(2.0f) there is the float logarithm. */
static double ln2f(void) { return (double)logf(2.0f); }

/* ------------------------------------------------------------------------ */
/* RLAX heat-flux relaxation of fluxes() — 3dmhdsub.h:322-442.               */
/* The reference runs it on nPEz z ranks (relax_nPEz here) of nRz planes     */
/* each; rank r's local plane l (0..nRz+1) is global plane r*nRz + l.  Each   */
/* rank scans CORRECT over its planes 1..nRz+1 (its upper ghost included),    */
/* with FCONM = 0 on that ghost (WW2 is never written on z ghost planes, so   */
/* the static array's zero stays there) and FRADM = the next rank's first     */
/* plane (received, :366-376) or, on the last rank, 0 (never received); the   */
/* ranks' end values are then added up (:383-421).  WW1 holds T' = TT - TTM   */
/* (:297-301), or under lShr the pressure RO*TT of :267 with TTM never        */
/* computed (zero); UU, VV, WW are the velocities of :274-277.                */

static void relax(orc_ctx* c)
{
    const mhd3d_desc* d = &c->d;
    const int zl = 1, zr = c->nz - 1;
    const int combo = (d->flag_iTCon == 0 && d->flag_izCon == 1) ? 1
                    : (d->flag_iTCon == 1 && d->flag_izCon == 0) ? 2 : 0;
    const double* TTM = c->TTM;
    double *FCONM = c->FCONM, *FRADM = c->FRADM, *CORRECT = c->CORRECT;
    const rng_t R = R_in(c);
    if ((c->TIMT > d->TSTART) && (c->TFAC > (double)1e-4f)) {               /* :323 */
        const double step = c->DT / ((MUT(27) ? 1.0 : 3.0) * d->TOFF);
        if (combo == 1) {                                                   /* :324-334 */
            if (TTM[zr - 1] <= c->TTM_MAX) {
                c->TFAC -= step;
            } else {
                c->TTM_MAX = TTM[zr - 1];
                c->TFAC += step;
                if (2.0 < c->TFAC) c->TFAC = 2.0;                           /* min(TFAC, 2.0f) */
            }
        } else if (combo == 2) {                                            /* :336-346 */
            if (TTM[zl] >= c->TTM_MIN) {
                c->TFAC -= step;
            } else {
                c->TTM_MIN = TTM[zl];
                c->TFAC += step;
                if (2.0 < c->TFAC) c->TFAC = 2.0;
            }
        } else {
            c->relax_exit = 2;                                              /* :347-348 */
            return;
        }
        f_mul2(c, SET, R, c->WW2, c->RW, MUT(23) ? c->TT : c->WW1, -2.5);   /* :349-353 */
        f_mul3(c, ADD, R, c->WW2, c->RW, c->UU, c->UU, -0.5);
        f_mul3(c, ADD, R, c->WW2, c->RW, c->VV, c->VV, -0.5);
        f_mul3(c, ADD, R, c->WW2, c->RW, c->WW, c->WW, -0.5);
        hmean(c, FCONM, c->WW2);
        const double hz = d->hz, REPR = d->REPR;
        const double *dzz = c->dzzdz, *RK = c->RKAPA;
        for (int z = zl; z < zr; z++)                                       /* :356-358 */
            FRADM[z] = (TTM[z + 1] - TTM[z - 1]) * hz * dzz[z] * RK[z] / REPR;
        FRADM[zl] = (-3.0 * TTM[zl] + 4.0 * TTM[zl + 1] - TTM[zl + 2]) * hz * dzz[zl] * RK[zl] / REPR;
        FRADM[zr - 1] = (3.0 * TTM[zr - 1] - 4.0 * TTM[zr - 2] + TTM[zr - 3]) * hz * dzz[zr - 1] *
                        RK[zr - 1] / REPR;
        double FTOT;                                                        /* :377-382 */
        if (d->flag_lRem) {
            const float k = 8.07f * 0.4f;   /* a float product, then promoted */
            FTOT = (double)k * d->THETA;
        } else {
            FTOT = d->THETA / REPR;
        }
        const int P = d->relax_nPEz, nRz = d->nPz / P;
        double* loc = (double*)calloc((size_t)(nRz + 2) * 2, sizeof(double));
        double* endval = (double*)calloc((size_t)P, sizeof(double));
        if (!loc || !endval) { free(loc); free(endval); c->relax_exit = 3; return; }
        double* ALPHA = loc;
        double* CL = loc + nRz + 2;
        for (int r = 0; r < P; ++r) {
            const int g0 = r * nRz;
            for (int l = 1; l <= nRz + 1; ++l) {                            /* :383-384 */
                const int g = g0 + l;
                const double fc = (l <= nRz || MUT(25)) ? FCONM[g] : 0.0;
                const double fr = (l <= nRz || r < P - 1) ? FRADM[g] : 0.0;
                double a = 1.0 - (fc + fr) / FTOT;
                if (!MUT(24)) a = a / (1.0 + 4.0 * fabs(a));
                ALPHA[l] = a;
            }
            CL[1] = 0;                                                      /* :385-388 */
            for (int l = 2; l <= nRz + 1; ++l)
                CL[l] = CL[l - 1] + 0.5 * (ALPHA[l] + ALPHA[l - 1]) * (c->ZEE[g0 + l] - c->ZEE[g0 + l - 1]);
            endval[r] = CL[nRz + 1];                                        /* :389-403 */
            for (int l = 1; l <= nRz; ++l) CORRECT[g0 + l] = CL[l];
        }
        for (int r = 0; r < P; ++r) {                                       /* :405-421 */
            double SUM = 0.0;
            if (combo == 1) {
                for (int k = 0; k < r; ++k) SUM += endval[k];
            } else {
                for (int k = P - 1; k >= r; --k) {
                    SUM -= endval[k];
                    if (!MUT(26)) SUM -= endval[k];
                }
            }
            for (int l = 1; l <= nRz; ++l) CORRECT[r * nRz + l] += SUM;
        }
        free(loc);
        free(endval);
    } else {
        for (int z = zl; z < zr; z++) CORRECT[z] = 0;                       /* :423-425 */
    }
    f_base_ext(c, ADD, R, c->FT, CORRECT, d->RLAX * c->TFAC);              /* :426 */
    if (combo != 0 && fmod(c->TIMT, 25.0) <= c->DT && (c->ICALL % 3 == 0)) /* :427-439 */
        c->relax_exit = 1;
    c->ICALL++;                                                             /* :441 */
}

/* ------------------------------------------------------------------------ */
/* fluxes() — 3dmhdsub.h:194-696                                             */

void orc_fluxes(orc_ctx* c)
{
    const mhd3d_desc* d = &c->d;
    const rng_t R = R_in(c), RE = R_ext(c), REXY = R_exy(c), REX = R_ex(c), REY = R_ey(c);
    const int zl = 1, zr = c->nz - 1, yl = 1, yr = c->ny - 1, xl = 1, xr = c->nx - 1;
    const int shr = d->flag_lShr != 0;
    double tmp;
    double *RU = c->RU, *RV = c->RV, *RW = c->RW, *RO = c->RO, *TT = c->TT, *RI = c->RI;
    double *UU = c->UU, *VV = c->VV, *WW = c->WW;
    double *FU = c->FU, *FV = c->FV, *FW = c->FW, *FR = c->FR, *FT = c->FT;
    double *WW1 = c->WW1, *WW2 = c->WW2, *WW3 = c->WW3;
    const double ORE = d->ORE, OCV = d->OCV, C13 = d->C13, C43 = d->C43;

    /* continuity :205-227 */
    f_d1(c, SET, DX, R, FR, RU, -1);
    f_d1(c, SUB, DY, R, FR, RV, 1);
    f_d1(c, SET, DZ, R, WW1, RW, 1);
    {
        const double TMPZ = d->hz * c->dzzdz[MUT(21) ? zl + 1 : zl];    /* :212-217 */
        for (int y = yl; y < yr; y++)
            for (int x = xl; x < xr; x++)
                WW1[IX(zl, y, x)] = (4.0 * RW[IX(zl + 1, y, x)] - RW[IX(zl + 2, y, x)]) * TMPZ;
    }
    {
        const double TMPZ = d->hz * c->dzzdz[zr - 1];                   /* :220-225 */
        for (int y = yl; y < yr; y++)
            for (int x = xl; x < xr; x++)
                WW1[IX(zr - 1, y, x)] = (RW[IX(zr - 3, y, x)] - 4.0 * RW[IX(zr - 2, y, x)]) * TMPZ;
    }
    f_base(c, SUB, R, FR, WW1, 1);

    /* density diffusion :229-264 */
    if (d->flag_iD != 0 || d->flag_lRem) {
        const size_t pl = (size_t)c->nx * c->ny;
        if (!MUT(9)) {
            memcpy(&RO[IX(0, 0, 0)], &RO[IX(2, 0, 0)], pl * sizeof(double));               /* :230-234 */
            memcpy(&RO[IX(c->nz - 1, 0, 0)], &RO[IX(c->nz - 3, 0, 0)], pl * sizeof(double)); /* :235-239 */
        }
        hmean(c, c->ROM, RO);
        f_base(c, SET, RE, WW1, RO, 1);
        if (!MUT(10)) f_base_ext(c, SUB, RE, WW1, c->ROM, 1);
    }
    if (d->flag_iD != 0) {
        if (d->flag_iD == 1) {
            for (int z = 0; z < c->nz; z++) c->WWZ[z] = 1.0 / c->ROM[z];
        } else {
            const double SGM = 1.0;
            const double CLN = -4.0 * ln2f() / SGM / SGM;
            for (int z = 0; z < c->nz; z++)
                c->WWZ[z] = exp(CLN * sq(c->ZEE[z])) + exp(CLN * sq(c->ZEE[z] - d->zMax));
        }
        if (d->DH != 0.0) {
            f_d2(c, SET, DX, R, WW2, WW1, 1);
            f_d2(c, SET, DY, R, WW3, WW1, 1);
            f_mul_ext(c, ADD, R, FR, WW2, c->WWZ, d->DH * ORE);
            f_mul_ext(c, ADD, R, FR, WW3, c->WWZ, d->DH * ORE);
        }
        if (d->DV != 0.0) {
            f_d2(c, SET, DZ, R, WW2, WW1, 1);
            f_mul_ext(c, ADD, R, FR, WW2, c->WWZ, d->DV * ORE);
        }
    }

    /* pressure, gravity :266-272 */
    f_base(c, SET, R, FW, RO, d->GRAV);
    f_mul2(c, SET, RE, WW1, RO, TT, 1);
    LOOP(RE, RI[i] = 1.0 / RO[i]);
    f_d1(c, SUB, DZ, R, FW, WW1, 1);
    f_d1(c, SET, DY, R, FV, WW1, -1);
    f_d1(c, SET, DX, R, FU, WW1, -1);

    /* velocities and advection :274-289 */
    f_mul2(c, SET, RE, UU, RU, RI, 1);
    f_mul2(c, SET, RE, VV, RV, RI, 1);
    f_mul2(c, SET, RE, WW, RW, RI, 1);
    f_mul_d1(c, SUB, DX, R, FU, RU, UU);
    f_mul_d1(c, SUB, DY, R, FU, RU, VV);
    f_mul_d1(c, SUB, DZ, R, FU, RU, WW);
    f_mul_d1(c, SUB, DX, R, FV, RV, UU);
    f_mul_d1(c, SUB, DY, R, FV, RV, VV);
    f_mul_d1(c, SUB, DZ, R, FV, RV, WW);
    f_mul_d1(c, SUB, DX, R, FW, RW, UU);
    f_mul_d1(c, SUB, DY, R, FW, RW, VV);
    f_mul_d1(c, SUB, DZ, R, FW, RW, WW);

    /* temperature advection and conduction :293-320 */
    if (!shr || MUT(12)) {
        f_d1_mul(c, SET, DX, R, FT, TT, UU, -1);
        f_d1_mul(c, SUB, DY, R, FT, TT, VV, 1);
        f_d1_mul(c, SUB, DZ, R, FT, TT, WW, 1);
    }
    if ((!shr || MUT(22)) && (d->RLAX > 0.0 || d->flag_lRem)) {             /* :297-302 */
        hmean(c, c->TTM, TT);
        LOOP(RE, WW1[i] = TT[i] - c->TTM[z]);
    }
    tmp = OCV / d->REPR;
    if (d->flag_lRem) {                                                     /* :304-313 */
        f_d2_mul(c, ADD, DX, R, FT, WW1, RI, tmp);
        f_d2_mul(c, ADD, DY, R, FT, WW1, RI, tmp);
        f_d2_mul(c, ADD, DZ, R, FT, WW1, RI, tmp);
        const double *TTM = c->TTM, *dzz = c->dzzdz;
        for (int z = zl; z < zr; z++) {
            c->HRAD[z] = (((TTM[z + 1] - 2.0 * TTM[z] + TTM[z - 1])) * d->h2z * dzz[z] * dzz[z] +
                          (TTM[z + 1] - TTM[z - 1]) * d->hz * c->d2zzdz2[z]) * c->RKAPA[z];
            if (!MUT(11)) c->HRAD[z] += (TTM[z + 1] - TTM[z - 1]) * d->hz * dzz[z] * c->DKAPA[z];
        }
        f_mul_ext(c, ADD, R, FT, RI, c->HRAD, tmp);
    } else {                                                           /* :315-319 */
        f_d2_mul_az(c, ADD, DX, R, FT, TT, RI, c->RKAPA, tmp);
        f_d2_mul_az(c, ADD, DY, R, FT, TT, RI, c->RKAPA, tmp);
        f_d2_mul_az(c, ADD, DZ, R, FT, TT, RI, c->RKAPA, tmp);
        f_d1_mul_az(c, ADD, DZ, R, FT, TT, RI, c->DKAPA, tmp);
    }
    if (d->RLAX > 0.0) relax(c);                                            /* :322-442 */
    if (d->flag_iTCon == 2) {                                               /* :443-461 */
        const double HH = d->HH, CLN = -4.0 * ln2f() / HH / HH;
        const double *EXX = c->EXX, *WYY = c->WYY, *ZEE = c->ZEE;
        if (d->TC != 0.0) {
            const double TIMT = c->TIMT;
            LOOP(R, FT[i] -= d->TP * OCV * 0.5 * (1.0 + tanh(ln2f() / d->QFH * (TIMT - d->TC))) * (MUT(13) ? RO[i] : RI[i]) *
                             exp(CLN * sq(EXX[x] - d->XP) / HH) * exp(CLN * sq(WYY[y] - d->YP) / HH) *
                             exp(CLN * sq(ZEE[z] - d->ZP) / HH));
        } else {
            LOOP(R, FT[i] -= d->TP * OCV * RI[i] * exp(CLN * sq(EXX[x] - d->XP) / HH) *
                             exp(CLN * sq(WYY[y] - d->YP) / HH) * exp(CLN * sq(ZEE[z] - d->ZP) / HH));
        }
    }

    /* viscous :462-472 */
    f_d2(c, ADD, DX, R, FU, UU, ORE * C43);
    f_d2(c, ADD, DY, R, FU, UU, ORE);
    f_d2(c, ADD, DZ, R, FU, UU, ORE);
    f_d2(c, ADD, DX, R, FV, VV, ORE);
    f_d2(c, ADD, DY, R, FV, VV, ORE * C43);
    f_d2(c, ADD, DZ, R, FV, VV, ORE);
    f_d2(c, ADD, DX, R, FW, WW, ORE);
    f_d2(c, ADD, DY, R, FW, WW, ORE);
    f_d2(c, ADD, DZ, R, FW, WW, ORE * C43);

    /* cross-viscous, viscous heating, pressure work :475-538 */
    f_d1(c, SET, DX, REY, WW1, UU, 1);
    if (MUT(5)) f_d1(c, SET, DY, R, WW2, VV, 1); else f_d1(c, SET, DX, REY, WW2, VV, 1);
    if (!shr) {
        f_d1(c, SET, DX, R, WW3, WW, 1);
        tmp = OCV * ORE;
        f_mul3(c, ADD, R, FT, RI, WW1, WW1, tmp * C43);
        f_mul3(c, ADD, R, FT, RI, WW2, WW2, tmp);
        if (!MUT(4)) f_mul3(c, ADD, R, FT, RI, WW3, WW3, tmp);
        f_mul2(c, SUB, R, FT, WW1, TT, OCV);
    }
    f_d1(c, ADD, DY, R, FU, WW2, C13 * ORE);
    f_d1(c, ADD, DY, R, FV, WW1, C13 * ORE);
    if (!shr) {
        f_d1(c, SET, DY, R, WW1, UU, 1);
        tmp = OCV * ORE;
        f_mul3(c, ADD, R, FT, RI, WW1, WW1, tmp);
        f_mul3(c, ADD, R, FT, RI, WW1, WW2, tmp * 2);
        f_d1(c, SET, DY, R, WW2, VV, 1);
        tmp = C43 * OCV * ORE;
        f_mul3(c, ADD, R, FT, RI, WW2, WW2, tmp);
        f_mul2(c, SUB, R, FT, WW2, TT, OCV);
        f_d1(c, SET, DX, R, WW1, UU, 1);
    }
    f_d1(c, SET, DZ, REXY, WW3, WW, 1);
    if (!shr) {
        tmp = C43 * OCV * ORE;
        f_mul3(c, ADD, R, FT, RI, WW3, WW3, tmp);
        f_mul3(c, ADD, R, FT, RI, WW1, WW3, -tmp);
        f_mul3(c, ADD, R, FT, RI, WW1, WW2, -tmp);
        f_mul3(c, ADD, R, FT, RI, WW2, WW3, -tmp);
        f_mul2(c, SUB, R, FT, WW3, TT, OCV);
    }
    f_d1(c, ADD, DX, R, FU, WW3, C13 * ORE);
    f_d1(c, ADD, DY, R, FV, WW3, C13 * ORE);
    f_d1(c, SET, DZ, REX, WW1, UU, 1);
    f_d1(c, SET, DZ, REY, WW2, VV, 1);
    if (!shr) {
        f_d1(c, SET, DY, R, WW3, WW, 1);
        tmp = OCV * ORE;
        f_mul3(c, ADD, R, FT, RI, WW1, WW1, tmp);
        f_mul3(c, ADD, R, FT, RI, WW2, WW2, tmp);
        f_mul3(c, ADD, R, FT, RI, WW3, WW3, tmp);
        f_mul3(c, ADD, R, FT, RI, WW2, WW3, 2.0 * tmp);
    }
    f_d1(c, ADD, DX, R, FW, WW1, C13 * ORE);
    f_d1(c, ADD, DY, R, FW, WW2, C13 * ORE);
    if (!shr) {
        f_d1(c, SET, DX, R, WW3, WW, 1);
        tmp = OCV * ORE;
        f_mul3(c, ADD, R, FT, RI, WW1, WW3, 2.0 * tmp);
    }

    /* Coriolis :540-545 */
    if (d->flag_lRot) {
        f_base(c, ADD, R, FU, RV, d->OMZ);
        f_base(c, SUB, R, FV, RU, d->OMZ);
        f_base(c, ADD, R, FV, RW, d->OMX);
        f_base(c, SUB, R, FW, RV, d->OMX);
    }

    /* induction, Lorentz force, Joule heating :547-695 */
    if (d->flag_lMag) {
        double *BX = c->BX, *BY = c->BY, *BZ = c->BZ;
        double *S1 = c->S1, *S2 = c->S2, *S3 = c->S3, *S4 = c->S4;
        const double OBETA = d->OBETA, ORM = d->ORM;
        const double KJ = OBETA * ORM * OCV, KJ2 = -2.0 * OBETA * ORM * OCV;

        f_d1(c, SET, DX, R, S1, UU, 1);
        f_d1(c, SET, DX, R, S2, VV, 1);
        f_d1(c, SET, DX, R, S3, WW, 1);
        f_mul2(c, SET, R, WW2, BX, S2, 1);
        f_mul2(c, SUB, R, WW2, BY, S1, 1);
        f_mul2(c, SET, R, WW3, BX, S3, 1);
        f_mul2(c, SUB, R, WW3, BZ, S1, 1);
        f_d1(c, SET, DY, R, S1, UU, 1);
        f_d1(c, SET, DY, R, S2, VV, 1);
        f_d1(c, SET, DY, R, S3, WW, 1);
        f_mul2(c, SET, R, WW1, BY, S1, 1);
        f_mul2(c, SUB, R, WW1, BX, S2, 1);
        f_mul2(c, MUT(3) ? SUB : ADD, R, WW3, BY, S3, 1);
        f_mul2(c, SUB, R, WW3, BZ, S2, 1);
        f_d1(c, SET, DZ, R, S1, UU, 1);
        f_d1(c, SET, DZ, R, S2, VV, 1);
        f_d1(c, SET, DZ, R, S3, WW, 1);
        for (int y = yl; y < yr; y++)                                   /* :572-577 */
            for (int x = xl; x < xr; x++)
                S3[IX(zl, y, x)] = (4.0 * WW[IX(zl + 1, y, x)] - WW[IX(zl + 2, y, x)]) * d->hz * c->dzzdz[zl];
        for (int y = yl; y < yr; y++)                                   /* :578-583 */
            for (int x = xl; x < xr; x++)
                S3[IX(zr - 1, y, x)] = (-4.0 * WW[IX(zr - 2, y, x)] + WW[IX(zr - 3, y, x)]) * d->hz * c->dzzdz[zr - 1];
        f_mul2(c, ADD, R, WW1, BZ, S1, 1);
        f_mul2(c, SUB, R, WW1, BX, S3, 1);
        f_mul2(c, ADD, R, WW2, BZ, S2, 1);
        f_mul2(c, SUB, R, WW2, BY, S3, 1);
        f_d1(c, SET, DX, R, S1, BX, 1);
        f_d1(c, SET, DX, R, S2, BY, 1);
        f_d1(c, SET, DX, R, S3, BZ, 1);
        f_d1(c, SET, DY, R, S4, BX, 1);
        f_mul2(c, ADD, R, FU, MUT(2) ? BZ : BY, S4, OBETA);
        f_mul2(c, SUB, R, FU, BY, S2, OBETA);
        f_mul2(c, SUB, R, FU, BZ, S3, OBETA);
        f_mul2(c, ADD, R, FV, BX, S2, OBETA);
        f_mul2(c, SUB, R, FV, BX, S4, OBETA);
        f_mul2(c, ADD, R, FW, BX, S3, OBETA);
        f_mul3(c, ADD, R, FT, RI, S4, S4, KJ);
        f_mul3(c, ADD, R, FT, RI, S2, S2, KJ);
        f_mul3(c, ADD, R, FT, RI, S3, S3, KJ);
        f_mul3(c, ADD, R, FT, RI, S2, S4, MUT(1) ? -KJ2 : KJ2);
        f_mul2(c, SUB, R, WW1, UU, S1, 1);
        f_mul2(c, SUB, R, WW1, VV, S4, 1);
        f_mul2(c, SUB, R, WW2, UU, S2, 1);
        f_mul2(c, SUB, R, WW3, UU, S3, 1);
        f_d2(c, SET, DX, R, S1, BX, 1);
        f_d2(c, SET, DX, R, S2, BY, 1);
        f_d2(c, SET, DY, R, S4, BX, 1);
        f_base(c, ADD, R, WW1, S1, ORM);
        f_base(c, ADD, R, WW1, S4, ORM);
        f_base(c, ADD, R, WW2, S2, ORM);
        f_d2(c, SET, DX, R, S4, BZ, 1);
        f_base(c, ADD, R, WW3, S4, ORM);
        f_d1(c, SET, DZ, R, S1, BX, 1);
        f_d1(c, SET, DY, R, S2, BY, 1);
        f_mul2(c, ADD, R, FU, BZ, S1, OBETA);
        f_mul2(c, SUB, R, FW, BX, S1, OBETA);
        f_mul3(c, ADD, R, FT, RI, S1, S1, KJ);
        f_mul3(c, ADD, R, FT, RI, S1, S3, KJ2);
        f_mul2(c, SUB, R, WW1, WW, S1, 1);
        f_mul2(c, SUB, R, WW2, VV, S2, 1);
        f_d2(c, SET, DZ, R, S1, BX, 1);
        {
            const double hz = d->hz, h2z = d->h2z;                      /* :640-654 */
            const double d2t = MUT(17) ? c->dzzdz[zl] : c->d2zzdz2[zl], dt_ = c->dzzdz[zl];
            const double d2b = c->d2zzdz2[zr - 1], db = c->dzzdz[zr - 1];
            for (int y = yl; y < yr; y++)
                for (int x = xl; x < xr; x++) {
                    const double b0 = BX[IX(zl, y, x)], b1 = BX[IX(zl + 1, y, x)];
                    const double b2 = BX[IX(zl + 2, y, x)], b3 = BX[IX(zl + 3, y, x)];
                    S1[IX(zl, y, x)] = (-3.0 * b0 + 4.0 * b1 - b2) * hz * d2t +
                                       (2.0 * b0 - 5.0 * b1 + 4.0 * b2 - b3) * h2z * dt_ * dt_;
                    const double e0 = BX[IX(zr - 1, y, x)], e1 = BX[IX(zr - 2, y, x)];
                    const double e2 = BX[IX(zr - 3, y, x)], e3 = BX[IX(MUT(6) ? zr - 3 : zr - 4, y, x)];
                    S1[IX(zr - 1, y, x)] = (3.0 * e0 - 4.0 * e1 + e2) * hz * d2b +
                                           (2.0 * e0 - 5.0 * e1 + 4.0 * e2 - e3) * h2z * db * db;
                }
        }
        f_d2(c, SET, DY, R, S2, BY, 1);
        f_base(c, ADD, R, WW1, S1, ORM);
        f_base(c, ADD, R, WW2, S2, ORM);
        f_d1(c, SET, DZ, R, S2, BY, 1);
        f_d1(c, SET, DY, R, S3, BZ, 1);
        f_base(c, SET, R, S1, S2, 1);
        f_base(c, SUB, R, S1, S3, 1);
        f_mul2(c, ADD, R, FV, BZ, S1, OBETA);
        f_mul2(c, SUB, R, FW, BY, S1, OBETA);
        f_mul3(c, ADD, R, FT, RI, S2, S2, KJ);
        f_mul3(c, ADD, R, FT, RI, S3, S3, KJ);
        f_mul3(c, ADD, R, FT, RI, S2, S3, KJ2);
        f_mul2(c, SUB, R, WW2, WW, S2, 1);
        f_mul2(c, SUB, R, WW3, VV, S3, 1);
        f_d2(c, SET, DZ, R, S2, BY, 1);
        {
            const double hz = d->hz, h2z = d->h2z;                      /* :673-687 */
            const double d2t = c->d2zzdz2[zl], dt_ = c->dzzdz[zl];
            const double d2b = c->d2zzdz2[zr - 1], db = c->dzzdz[zr - 1];
            for (int y = yl; y < yr; y++)
                for (int x = xl; x < xr; x++) {
                    const double b0 = BY[IX(zl, y, x)], b1 = BY[IX(zl + 1, y, x)];
                    const double b2 = BY[IX(zl + 2, y, x)], b3 = BY[IX(zl + 3, y, x)];
                    S2[IX(zl, y, x)] = (-3.0 * b0 + 4.0 * b1 - b2) * hz * d2t +
                                       (2.0 * b0 - 5.0 * b1 + 4.0 * b2 - b3) * h2z * dt_ * dt_;
                    const double e0 = BY[IX(zr - 1, y, x)], e1 = BY[IX(zr - 2, y, x)];
                    const double e2 = BY[IX(zr - 3, y, x)], e3 = BY[IX(zr - 4, y, x)];
                    S2[IX(zr - 1, y, x)] = (3.0 * e0 - 4.0 * e1 + e2) * hz * d2b +
                                           (2.0 * e0 - 5.0 * e1 + 4.0 * e2 - e3) * h2z * db * db;
                }
        }
        f_d2(c, SET, DY, R, S3, BZ, 1);
        f_base(c, ADD, R, WW2, S2, ORM);
        f_base(c, ADD, R, WW3, S3, ORM);
        f_d1(c, SET, DZ, R, S3, BZ, 1);
        f_mul2(c, SUB, R, WW3, WW, S3, 1);
        f_d2(c, SET, DZ, R, S3, BZ, 1);
        f_base(c, ADD, R, WW3, S3, ORM);
    }
}

/* ------------------------------------------------------------------------ */
/* BCON() — 3dmhdsub.h:720-873 (iTCon 0-4, the TP plume of iTCon 0/1, izCon  */
/* 0/1, lRem and lShr walls, iBCon 0)                                         */

void orc_bcon(orc_ctx* c)
{
    const mhd3d_desc* d = &c->d;
    const int zl = 1, zr = c->nz - 1, yl = 1, yr = c->ny - 1, xl = 1, xr = c->nx - 1;
    const int nx = c->nx;
    double *RU = c->RU, *RV = c->RV, *RW = c->RW, *RO = c->RO, *TT = c->TT, *BZ = c->BZ;
    const double C13 = d->C13, C23 = d->C23, C43 = d->C43;
    const double RRT = 1.0, RRB = -1.0, RPP = 0.0;

    /* top wall, myPEz == 0 */
    for (int y = yl; y < yr; y++)
        for (int x = xl; x < xr; x++) {
            if (d->flag_lShr) {
                RU[IX(zl, y, x)] = 0.0;
                RV[IX(zl, y, x)] = RRT * cos(RPP * c->TIMT) * RO[IX(zl, y, x)];
            } else {
                RU[IX(zl, y, x)] = RO[IX(zl, y, x)] * (C43 * RU[IX(zl + 1, y, x)] / RO[IX(zl + 1, y, x)] -
                                                       C13 * RU[IX(zl + 2, y, x)] / RO[IX(zl + 2, y, x)]);
                RV[IX(zl, y, x)] = RO[IX(zl, y, x)] * (C43 * RV[IX(zl + 1, y, x)] / RO[IX(zl + 1, y, x)] -
                                                       C13 * RV[IX(zl + 2, y, x)] / RO[IX(zl + 2, y, x)]);
            }
        }
    for (int y = yl; y < yr; y++)
        for (int x = xl; x < xr; x++) RW[IX(zl, y, x)] = 0.0;
    if (d->TP != 0.0 && (d->flag_iTCon == 0 || d->flag_iTCon == 1)) {             /* :750-776 */
        const double TPR = d->TC != 0.0 ? d->TP + (1.0 - d->TP) * exp(-c->TIMT / d->TC) : d->TP;
        const double HH = d->HH, CLN = -4.0 * ln2f() / HH / HH;
        if (d->flag_iTCon == 0) {
            for (int y = yl; y < yr; y++)
                for (int x = xl; x < xr; x++)
                    TT[IX(zl, y, x)] = 1.0 - (1.0 - TPR) * exp(CLN * sq(c->EXX[x] - d->XP) / (MUT(14) ? HH : 1.0)) / HH *
                                                 exp(CLN * sq(c->WYY[y] - d->YP)) / HH;
        } else {
            const double DZ = d->ORZD / c->dzzdz[zl];
            for (int y = yl; y < yr; y++)
                for (int x = xl; x < xr; x++)
                    TT[IX(zl, y, x)] = C43 * TT[IX(zl + 1, y, x)] - C13 * TT[IX(zl + 2, y, x)] -
                                       C23 * DZ * (d->THETA - (d->THETA - TPR) * exp(CLN * sq(c->EXX[x] - d->XP)) / HH *
                                                                  exp(CLN * sq(c->WYY[y] - d->YP)) / HH);
        }
    } else if (d->flag_iTCon == 0 || d->flag_iTCon == 2 || d->flag_iTCon == 4) {      /* :777-795 */
        for (int y = yl; y < yr; y++)
            for (int x = xl; x < xr; x++) {
                if (d->flag_lRem) {
                    TT[IX(zl, y, x)] = d->TU;
                } else {
                    TT[IX(zl, y, x)] = 1.0;
                    if (d->flag_iTCon == 4 && x < nx / 2) TT[IX(zl, y, x)] = 1.0 - d->TP;
                }
            }
    } else if (d->flag_iTCon == 1) {                                          /* :796-809 */
        const double DZ = d->ORZF / c->dzzdz[zl];
        for (int y = yl; y < yr; y++)
            for (int x = xl; x < xr; x++) {
                if (d->flag_lRem)
                    TT[IX(zl, y, x)] = C43 * TT[IX(zl + 1, y, x)] - C13 * TT[IX(zl + 2, y, x)] + d->DZTU;
                else
                    TT[IX(zl, y, x)] = C43 * TT[IX(zl + 1, y, x)] - C13 * TT[IX(zl + 2, y, x)] -
                                       C23 * DZ * d->THETA;
            }
    }
    if (d->flag_lMag && d->flag_iBCon == 0)
        for (int y = yl; y < yr; y++)
            for (int x = xl; x < xr; x++) BZ[IX(zl, y, x)] = 0.0;

    /* bottom wall, myPEz == nPEz-1 */
    for (int y = yl; y < yr; y++)
        for (int x = xl; x < xr; x++) {
            if (d->flag_lShr) {
                RU[IX(zr - 1, y, x)] = 0.0;
                RV[IX(zr - 1, y, x)] = RRB * cos(RPP * c->TIMT) * RO[IX(zr - 1, y, x)];
            } else {
                RU[IX(zr - 1, y, x)] = RO[IX(zr - 1, y, x)] * (C43 * RU[IX(zr - 2, y, x)] / RO[IX(zr - 2, y, x)] -
                                                               C13 * RU[IX(zr - 3, y, x)] / RO[IX(zr - 3, y, x)]);
                RV[IX(zr - 1, y, x)] = RO[IX(zr - 1, y, x)] * (C43 * RV[IX(zr - 2, y, x)] / RO[IX(zr - 2, y, x)] -
                                                               C13 * RV[IX(zr - 3, y, x)] / RO[IX(zr - 3, y, x)]);
            }
        }
    for (int y = yl; y < yr; y++)
        for (int x = xl; x < xr; x++) RW[IX(zr - 1, y, x)] = 0.0;
    if (d->flag_izCon == 0) {                                                 /* :842-847 */
        for (int y = yl; y < yr; y++)
            for (int x = xl; x < xr; x++) TT[IX(zr - 1, y, x)] = d->TB;
    } else if (d->flag_lRem) {                                                /* :850-854 */
        for (int y = yl; y < yr; y++)
            for (int x = xl; x < xr; x++)
                TT[IX(zr - 1, y, x)] = C43 * TT[IX(zr - 2, y, x)] - C13 * TT[IX(zr - 3, y, x)] + d->DZTB;
    } else {                                                             /* :848-861 */
        const double DZ = d->ORZF / c->dzzdz[zr - 1];
        for (int y = yl; y < yr; y++)
            for (int x = xl; x < xr; x++)
                TT[IX(zr - 1, y, x)] = C43 * TT[IX(zr - 2, y, x)] - (MUT(7) ? C23 : C13) * TT[IX(zr - 3, y, x)] -
                                       C23 * DZ * d->THETA / c->RKAPA[zr - 1];
    }
    if (d->flag_lMag && d->flag_iBCon == 0)
        for (int y = yl; y < yr; y++)
            for (int x = xl; x < xr; x++) BZ[IX(zr - 1, y, x)] = 0.0;
}

/* ------------------------------------------------------------------------ */
/* TSTEP — 3dmhdsub.h:36-106 (ISW == 1 branch)                               */

static inline double smax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static inline double smin(double a, double b) { return (b < a) ? b : a; } /* std::min */

double orc_cfl(orc_ctx* c, double wmin[4], double* umach)
{
    const mhd3d_desc* d = &c->d;
    const rng_t R = R_in(c);
    double *RU = c->RU, *RV = c->RV, *RW = c->RW, *RO = c->RO, *TT = c->TT, *UU = c->UU;
    double *BX = c->BX, *BY = c->BY, *BZ = c->BZ;
    double *WW1 = c->WW1, *WW2 = c->WW2, *WW3 = c->WW3, *VV = c->VV;
    const double OGAMMA = 1.0 / d->GAMMA, GAMMA = d->GAMMA, OBETA = d->OBETA;
    const int xdim = c->nx > 3;

    LOOP(R, UU[i] = (RU[i] * RU[i] + RV[i] * RV[i] + RW[i] * RW[i]) / (RO[i] * RO[i]));
    double um = 0.0;
    for (int z = R.zl; z < R.zr; ++z)
        for (int y = R.yl; y < R.yr; ++y)
            for (int x = R.xl; x < R.xr; ++x) {
                const size_t i = IX(z, y, x);
                um = smax(um, OGAMMA * UU[i] / TT[i]);
            }
    c->UMACH = sqrt(um);
    if (d->flag_lMag) {
        LOOP(R, UU[i] += GAMMA * TT[i] + (BX[i] * BX[i] + BY[i] * BY[i] + BZ[i] * BZ[i]) * OBETA / RO[i] +
                         2.0 * sqrt(UU[i] * (GAMMA * TT[i] + (BX[i] * BX[i] + BY[i] * BY[i] + BZ[i] * BZ[i]) * OBETA / RO[i])));
    } else {
        LOOP(R, UU[i] += (MUT(8) ? 1.0 : GAMMA) * TT[i] + 2.0 * sqrt(GAMMA * UU[i] * TT[i]));
    }
    const double *DDX = c->DDX, *DDY = c->DDY, *DDZ = c->DDZ, *RKAPA = c->RKAPA;
    const double REPR = d->REPR, CV = d->CV, RE = d->RE, RM = d->RM;
    LOOP(R, {
        double DD = smin(DDY[y], DDZ[z]);
        if (xdim) DD = smin(DD, DDX[x]);
        WW1[i] = DD / sqrt(UU[i]);
        if (d->flag_lRem)                                                    /* :78-82 */
            WW2[i] = 0.5 * DD * DD * REPR * CV * RO[i] / ((MUT(15) ? 0.0 : 1.0) + RKAPA[z]);
        else
            WW2[i] = 0.5 * DD * DD * REPR * CV * RO[i] / RKAPA[z];
        WW3[i] = 0.375 * DD * DD * RE * RO[i];
        if (d->flag_lMag) VV[i] = 0.5 * DD * DD * (MUT(18) ? RE : RM);
    });
    double w[4] = {1e99, 1e99, 1e99, 1e99};
    for (int z = R.zl; z < R.zr; ++z)
        for (int y = R.yl; y < R.yr; ++y)
            for (int x = R.xl; x < R.xr; ++x) {
                const size_t i = IX(z, y, x);
                w[0] = smin(w[0], WW1[i]);
                w[1] = smin(w[1], WW2[i]);
                w[2] = smin(w[2], WW3[i]);
                if (d->flag_lMag) w[3] = smin(w[3], VV[i]);
            }
    double DT;
    if (d->flag_lMag)
        DT = d->SFF * smin(smin(w[0], w[1]), smin(w[2], w[3]));
    else
        DT = d->SFF * smin(smin(w[0], w[1]), w[2]);
    if (wmin) memcpy(wmin, w, sizeof(w));
    if (umach) *umach = c->UMACH;
    return DT;
}

/* ------------------------------------------------------------------------ */
/* step() — 3dmhdsub.h:31-153                                                */

void orc_stage(orc_ctx* c, int k, double DT)
{
    const mhd3d_desc* d = &c->d;
    c->DT = DT;   /* the global DT that fluxes() reads (RLAX) */
    const rng_t R = R_in(c);
    const int mag = d->flag_lMag != 0;
    if (k == 1) {                                                      /* :155-166 */
        f_copy(c, R, c->ZRU, c->RU);
        f_copy(c, R, c->ZRV, c->RV);
        f_copy(c, R, c->ZRW, c->RW);
        f_copy(c, R, c->ZRO, c->RO);
        f_copy(c, R, c->ZTT, c->TT);
        if (mag) {
            f_copy(c, R, c->ZBX, c->BX);
            f_copy(c, R, c->ZBY, c->BY);
            f_copy(c, R, c->ZBZ, c->BZ);
        }
    } else {                                                           /* :168-179 */
        const double COEF = (k == 2 ? d->ZETA1 : d->ZETA2) * DT;
        f_scale_add(c, R, c->ZRU, c->RU, c->FU, COEF);
        f_scale_add(c, R, c->ZRV, c->RV, c->FV, COEF);
        f_scale_add(c, R, c->ZRW, c->RW, c->FW, COEF);
        f_scale_add(c, R, c->ZRO, c->RO, c->FR, COEF);
        f_scale_add(c, R, c->ZTT, c->TT, c->FT, COEF);
        if (mag) {
            f_scale_add(c, R, c->ZBX, c->BX, c->WW1, COEF);
            f_scale_add(c, R, c->ZBY, c->BY, c->WW2, COEF);
            f_scale_add(c, R, c->ZBZ, c->BZ, c->WW3, COEF);
        }
    }
    orc_fluxes(c);
    {                                                                  /* :181-192 */
        const double COEF = (k == 1 ? d->GAM1 : k == 2 ? d->GAM2 : d->GAM3) * DT;
        f_scale_add(c, R, c->RU, c->ZRU, c->FU, COEF);
        f_scale_add(c, R, c->RV, c->ZRV, c->FV, COEF);
        f_scale_add(c, R, c->RW, c->ZRW, c->FW, COEF);
        f_scale_add(c, R, c->RO, c->ZRO, c->FR, COEF);
        f_scale_add(c, R, c->TT, c->ZTT, c->FT, COEF);
        if (mag) {
            f_scale_add(c, R, c->BX, c->ZBX, c->WW1, COEF);
            f_scale_add(c, R, c->BY, c->ZBY, c->WW2, COEF);
            f_scale_add(c, R, c->BZ, c->ZBZ, c->WW3, COEF);
        }
    }
    orc_bcon(c);
    orc_communicate(c);
}

double orc_step(orc_ctx* c, int fixed_dt, double dt)
{
    if (fixed_dt)
        c->DT = dt;
    else
        c->DT = orc_cfl(c, NULL, NULL);
    orc_stage(c, 1, c->DT);
    orc_stage(c, 2, c->DT);
    orc_stage(c, 3, c->DT);
    c->nIt += 1;                                                       /* :150-152 */
    c->TIMC += c->DT;
    c->TIMT = c->TIMI + c->TIMC;
    return c->DT;
}

/* ------------------------------------------------------------------------ */
/* RAN2 — 3dmhdphsic.h:523-547.  state[0]=IDUM, state[1]=IIY, state[2..98]=IIR */

double orc_ran2(int* st)
{
    const int M = 714025, IA = 1366, IC = 150889;
    const double RM = (double)(1.0f / (float)M);
    int* IDUM = &st[0];
    int* IIY = &st[1];
    int* IIR = &st[2];
    if (*IDUM < 0) {
        *IDUM = (int)fmod((double)(IC - *IDUM), (double)M);
        for (int j = 0; j < 97; ++j) {
            *IDUM = (int)fmod((double)(IA * *IDUM + IC), (double)M);
            IIR[j] = *IDUM;
        }
        *IDUM = (int)fmod((double)(IA * *IDUM + IC), (double)M);
        *IIY = *IDUM;
    }
    int j = (97 * *IIY) / M;
    if (j >= 97 || j < 0) return -1.0;
    *IIY = IIR[j];
    double r = (double)(float)(*IIY) * RM;
    *IDUM = (IA * *IDUM + IC) % M;
    IIR[j] = *IDUM;
    return r;
}

void orc_static_rnd(int nPx, int nPy, int nzg, int nPEy, int nPEz, double AMPT, double* out)
{
    const int nRy = nPy / nPEy, nRz = nzg / nPEz;
    const size_t NXg = (size_t)nPx + 2, NYg = (size_t)nPy + 2;
    int st[99];
    memset(st, 0, sizeof(st));
    st[0] = -62659;                                                    /* :708 */
    for (int nnz = 0; nnz < nPEz; nnz++)
        for (int k = 1; k <= nRz; k++)
            for (int nny = 0; nny < nPEy; nny++)
                for (int y = 1; y <= nRy; y++)
                    for (int x = 1; x <= nPx; x++) {
                        const double v = 1.0 + AMPT * (orc_ran2(st) - 0.5);
                        const size_t gz = (size_t)nnz * nRz + k;
                        const size_t gy = (size_t)nny * nRy + y;
                        out[(gz * NYg + gy) * NXg + x] = v;
                    }
}

/* ------------------------------------------------------------------------ */
/* Bulirsch-Stoer polytrope of Static() (3dmhdphsic.h:230-355, 596-609)      */

typedef struct {
    double THETA, GRAV, zMax;
    double rz_X[11], rz_FX[7];          /* rzextr statics   */
    double mm_YM[10], mm_YN[10];        /* mmid statics     */
    double bs_YERR[10], bs_YSeq[10];    /* bsstep statics   */
} bs_state;

static void bs_derivs(bs_state* s, double zz, const double* Y, double* DYDX)
{
    /* mkGrid<Z> with nGrid == 0 (3dmhdphsic.h:37-41); PZP == 0 so RKAP = 1 */
    const double dzzdz = 1.0 / s->zMax;
    (void)zz;
    const double RKAP = 1.0;
    DYDX[0] = s->THETA / RKAP / dzzdz;
    DYDX[1] = s->GRAV / (1.0 + Y[0]) / dzzdz;
}

static void bs_mmid(bs_state* s, const double* Y, const double* DYDX, int NVAR, double XS,
                    double HTOT, int NSTP, double* YOUT)
{
    double* YM = s->mm_YM;
    double* YN = s->mm_YN;
    double H = HTOT / NSTP, X = XS + H;
    for (int i = 0; i < NVAR; ++i) YM[i] = Y[i], YN[i] = Y[i] + H * DYDX[i];
    bs_derivs(s, X, YN, YOUT);
    double H2 = 2.0 * H;
    for (int N = 2; N <= NSTP; ++N) {
        for (int i = 0; i < NVAR; ++i) {
            double SWAP = YM[i] + H2 * YOUT[i];
            YM[i] = YN[i];
            YN[i] = SWAP;
        }
        X = X + H;
        bs_derivs(s, X, YN, YOUT);
    }
    for (int i = 0; i < NVAR; ++i) YOUT[i] = 0.5 * (YM[i] + YN[i] + H * YOUT[i]);
}

static void bs_rzextr(bs_state* s, int IEST, double XEST, const double* YEST, double* YZ,
                      double D[10][7], double* DY, int NV, int NUSE)
{
    double* X = s->rz_X;
    double* FX = s->rz_FX;
    X[IEST] = XEST;
    if (IEST == 0) {
        for (int j = 0; j < NV; ++j) YZ[j] = D[j][0] = DY[j] = YEST[j];
    } else {
        int M1 = IEST + 1 < NUSE ? IEST + 1 : NUSE;
        for (int k = 1; k < M1; ++k) FX[k] = X[IEST - k] / XEST;
        for (int j = 0; j < NV; ++j) {
            double YY = YEST[j], V = D[j][0], C = YY, DDY = 0.0;
            D[j][0] = YY;
            for (int k = 1; k < M1; ++k) {
                double B1 = FX[k] * V;
                double B = B1 - C;
                if (B != 0.0) {
                    B = (C - V) / B;
                    DDY = C * B;
                    C = B1 * B;
                } else {
                    DDY = V;
                }
                V = D[j][k];
                D[j][k] = DDY;
                YY = YY + DDY;
            }
            DY[j] = DDY;
            YZ[j] = YY;
        }
    }
}

static int bs_step(bs_state* s, double* Y, const double* DYDX, int NV, double* X, double HTRY,
                   double EPS, double* HDID)
{
    const int IMAX = 11, NUSE = 7;
    static const int NSeq[11] = {2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96};
    double H = HTRY, XSAV = *X;
    double YSAV[10] = {0}, DYSAV[10] = {0};
    double D[10][7];
    memset(D, 0, sizeof(D));
    for (int i = 0; i < NV; ++i) YSAV[i] = Y[i], DYSAV[i] = DYDX[i];
    for (int guard = 0; guard < 64; ++guard) {
        for (int i = 0; i < IMAX; ++i) {
            bs_mmid(s, YSAV, DYSAV, NV, XSAV, H, NSeq[i], s->bs_YSeq);
            double XEST = (H / NSeq[i]) * (H / NSeq[i]);
            bs_rzextr(s, i, XEST, s->bs_YSeq, Y, D, s->bs_YERR, NV, NUSE);
            double ERRMAX = 0.0;
            for (int j = 0; j < NV; ++j) ERRMAX = smax(ERRMAX, fabs(s->bs_YERR[j] / Y[j]));
            ERRMAX /= EPS;
            if (ERRMAX < 1) {
                *X = *X + H;
                *HDID = H;
                return 0;
            }
        }
        H = 0.25 * H / pow(2.0, (IMAX - NUSE) / 2);
        if (*X + H == *X) return -1;
    }
    return -1;
}

int orc_static_profile(int nPz, double THETA, double GRAV, double zMax, double* T, double* R)
{
    bs_state s;
    memset(&s, 0, sizeof(s));
    s.THETA = THETA;
    s.GRAV = GRAV;
    s.zMax = zMax;
    const double EPS = (double)1.0E-12f;
    const double DZZ = (double)(1.0f / (float)(nPz - 1));
    double Y[2] = {0.0, 0.0}, DYDX[2];
    T[0] = 1.0, R[0] = 1.0;
    double ZZ = 0.0;
    for (int k = 1; k < nPz; ++k) {
        double HTRY = DZZ, HDID = 0.0;
        bs_derivs(&s, ZZ, Y, DYDX);
        if (bs_step(&s, Y, DYDX, 2, &ZZ, HTRY, EPS, &HDID) != 0) return -1;
        if (HDID != HTRY) return -1;
        T[k] = 1.0 + Y[0];
        R[k] = exp(Y[1]) / (1.0 + Y[0]);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* M. Rempel's background of Static() under lRem (3dmhdphsic.h:357-522,      */
/* 565-592): getbackground integrates derivs1 from x = 2 with (T, P) = (1, 1) */
/* down to each plane's height by ODEINT (RKQC step doubling of RK4).        */

static double rp_kappa(double z, double rho) { return rho + (1.25f - rho) * exp(-5.0f * z); }

static void rp_derivs(double x, const double* yt, double* dyt)
{
    const double nad = 0.41f, os = 1.0f;
    double krad = rp_kappa(x, yt[1] / yt[0]);
    double nrad = 0.4f / krad;
    if (nrad <= os * nad) {
        dyt[0] = nrad;
        dyt[1] = yt[1] / yt[0];
    } else {
        dyt[0] = nad;
        dyt[1] = yt[1] / yt[0];
    }
}

static void rp_rk4(const double* Y, const double* DYDX, double X, double H, double* YOUT)
{
    double YT[2], DYT[2], DYM[2];
    double HH = H * 0.5f, H6 = H / 6.0f, XH = X + HH;
    for (int i = 0; i < 2; i++) YT[i] = Y[i] + HH * DYDX[i];
    rp_derivs(XH, YT, DYT);
    for (int i = 0; i < 2; i++) YT[i] = Y[i] + HH * DYT[i];
    rp_derivs(XH, YT, DYM);
    for (int i = 0; i < 2; i++) {
        YT[i] = Y[i] + H * DYM[i];
        DYM[i] = DYT[i] + DYM[i];
    }
    rp_derivs(X + H, YT, DYT);
    for (int i = 0; i < 2; i++) YOUT[i] = Y[i] + H6 * (DYDX[i] + DYT[i] + 2.0f * DYM[i]);
}

static int rp_rkqc(double* Y, double* DYDX, double* X, double HTRY, double EPS, const double* YSCAL,
                   double* HDID, double* HNEXT)
{
    const double FCOR = 1.0f / 15.0f, SAFETY = 0.9f, ERRCON = 6.0e-4f;
    const double PGROW = -0.20f, PSHRNK = -0.25f;
    double YTEMP[2], YSAV[2], DYSAV[2];
    double XSAV = *X, H = HTRY;
    for (int i = 0; i < 2; i++) YSAV[i] = Y[i], DYSAV[i] = DYDX[i];
    for (;;) {
        double HH = 0.5f * H;
        rp_rk4(YSAV, DYSAV, XSAV, HH, YTEMP);
        *X = XSAV + HH;
        rp_derivs(*X, YTEMP, DYDX);
        rp_rk4(YTEMP, DYDX, *X, HH, Y);
        *X = XSAV + H;
        if (*X == XSAV) return -1;
        rp_rk4(YSAV, DYSAV, XSAV, H, YTEMP);
        double ERRMAX = 0.0f;
        for (int i = 0; i < 2; i++) {
            YTEMP[i] = Y[i] - YTEMP[i];
            ERRMAX = smax(ERRMAX, fabs(YTEMP[i] / YSCAL[i]));
        }
        ERRMAX = ERRMAX / EPS;
        if (ERRMAX > 1.0) {
            H = SAFETY * H * pow(ERRMAX, PSHRNK);
        } else {
            *HDID = H;
            *HNEXT = ERRMAX > ERRCON ? SAFETY * H * pow(ERRMAX, PGROW) : 4.0f * H;
            break;
        }
    }
    for (int i = 0; i < 2; i++) Y[i] += YTEMP[i] * FCOR;
    return 0;
}

static int rp_background(double x2, double* t, double* rho)
{
    const double TINY = 1.0e-30f, EPS = 1e-8, H1 = 1e-4, HMIN = 1e-50;
    const double X1 = 2.0f;
    double Y[2] = {1, 1}, DYDX[2], YSCAL[2];
    double X = X1, H = (x2 > X1) ? H1 : -H1;
    for (int n = 1; n <= 1000; ++n) {
        rp_derivs(X, Y, DYDX);
        for (int i = 0; i < 2; ++i) YSCAL[i] = fabs(Y[i]) + fabs(H * DYDX[i]) + TINY;
        if ((X + H - x2) * (X + H - X1) > 0) H = x2 - X;
        double HDID = 0, HNEXT = 0;
        if (rp_rkqc(Y, DYDX, &X, H, EPS, YSCAL, &HDID, &HNEXT)) return -1;
        if ((X - x2) * (x2 - X1) >= 0) {
            *t = Y[0];
            *rho = Y[1] / Y[0];
            return 0;
        }
        if (fabs(HNEXT) < HMIN) return -1;
        H = HNEXT;
    }
    return -1;
}

int orc_rempel_profile(int nPz, double THETA, double REPR, double zMax, double* T, double* R,
                       double* RK, double* DRK, double walls[3])
{
    if (nPz < 3) return -1;
    const double hz = (double)(0.5f * (float)(nPz - 1));
    for (int k = 0; k < nPz; k++) {
        double SZ = (double)((float)k / (float)(nPz - 1)) * zMax;     /* mkGrid, nGrid 0 */
        if (rp_background(SZ, &T[k], &R[k])) return -1;
        RK[k] = rp_kappa(SZ, R[k]);
        RK[k] *= 8.07f * THETA * REPR;
    }
    const double SDZZDZ = 1.0f / zMax;
    DRK[0] = (-3.0f * RK[0] + 4.0f * RK[1] - RK[2]) * hz * SDZZDZ;
    for (int k = 1; k < nPz - 1; ++k) DRK[k] = (RK[k + 1] - RK[k - 1]) * hz * SDZZDZ;
    DRK[nPz - 1] = (3.0f * RK[nPz - 1] - 4.0f * RK[nPz - 2] + RK[nPz - 3]) * hz * SDZZDZ;
    walls[0] = T[0];
    walls[1] = T[0] - 4.0f / 3.0f * T[1] + 1.0f / 3.0f * T[2];
    walls[2] = T[nPz - 1] - 4.0f / 3.0f * T[nPz - 2] + 1.0f / 3.0f * T[nPz - 3];
    return 0;
}
