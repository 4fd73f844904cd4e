// Device-side types of the MI355X RK3 timestep.
//
// Device layout of one register (U or Z): the fields RU,RV,RW,RO,TT[,BX,BY,BZ]
// one after another, each [nz][nPy][pitch] doubles with nz = nPz+2.  The z
// ghost planes of the reference's arrays (3dmhddata.h:8) are kept; the x and
// y ghost columns are not: x and y are periodic (3dmhdmpi.h:59-142), so the
// kernels wrap the in-plane indices instead of refilling ghosts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Linkage of the non-template kernels defined in these headers: external in
// the main translation unit, which launches them; internal (and so not
// emitted, being unused) in the translation units of the FP32 kernels and of
// the off-default interior kernels, which include the same headers.
#if defined(MHD3D_FP32_FORM_TU) || defined(MHD3D_SECONDARY_TU)
#define MHD_TU_LOCAL static
#else
#define MHD_TU_LOCAL
#endif

namespace mhd {

constexpr int NF_HYDRO = 5;
constexpr int NF_MAG = 8;

// Sizes of one domain.
struct Geom {
    int nPx, nPy, nPz;   // interior x, y; z planes incl. walls
    int nz;              // nPz + 2 (with the z ghost planes)
    int pitch;           // row pitch in doubles (>= nPx)
    int64_t plane;       // nPy * pitch
    int64_t fs;          // field stride in doubles
};

// Scalars of set() (3dmhd.cpp:353-495) the kernels read.
struct Phys {
    double C13, C23, C43;
    double GAM[3], ZETA[2];
    double SFF;
    double hx, h2x, hy, h2y, hz, h2z;
    double ORZF;
    double RE, ORE, REPR, CV, OCV, GAMMA, OGAMMA, THETA, GRAV;
    double OMX, OMZ, RM, ORM, OBETA;
    double TP, TB;
    // products of the scalars above that fluxes() forms before its sweeps
    // (3dmhdsub.h:315,462,479-536,604-668), evaluated once on the host with
    // the same double operations so the kernels read them as constants
    double ORE_C43;       // ORE * C43
    double C13_ORE;       // C13 * ORE
    double OCV_REPR;      // OCV / REPR
    double OCV_ORE;       // OCV * ORE
    double OCV_ORE_C43;   // (OCV * ORE) * C43
    double OCV_ORE_2;     // (OCV * ORE) * 2
    double C43_OCV_ORE;   // (C43 * OCV) * ORE
    double KJ;            // (OBETA * ORM) * OCV
    double KJ2;           // ((-2 * OBETA) * ORM) * OCV
    int izCon, iTCon, iBCon, xdim;
};

// TSTEP inputs of one plane z in one 128-byte record: its metric terms and
// the scalars of the limits (3dmhdsub.h:41-88).  The streaming kernel's
// stage 3 reads the record of the plane it updates, so these values are
// loaded where they are used instead of being held through the whole sweep.
struct CflRec {
    double DDZ, KCFL, OGAMMA, GAMMA, OBETA, REPR, CV, RE, RM;
    double pad[7];
};

// 1-D metric arrays on the device, indexed with the reference's ghosted
// indices (interior x = 1..nPx etc.).
struct Metric {
    const double* __restrict__ dxxdx;
    const double* __restrict__ d2xxdx2;
    const double* __restrict__ DDX;
    const double* __restrict__ dyydy;
    const double* __restrict__ d2yydy2;
    const double* __restrict__ DDY;
    const double* __restrict__ dzzdz;
    const double* __restrict__ d2zzdz2;
    const double* __restrict__ DDZ;
    const double* __restrict__ RKAPA;
    const double* __restrict__ DKAPA;
    // denominator of the TSTEP conduction limit WW2 per plane: RKAPA[z], or
    // 1.0f + RKAPA[z] with lRem (3dmhdsub.h:78-82), formed on the host
    const double* __restrict__ KCFL;
    const CflRec* __restrict__ cflz;   // one record per plane z
};

// Off-default branches of fluxes(), BCON() and TSTEP that the point kernel
// and the wall kernel evaluate (bits of ExtArgs::flags).
enum : int {
    EXT_DH = 1,      // iD != 0, DH != 0: horizontal density diffusion (3dmhdsub.h:255-260)
    EXT_DV = 2,      // iD != 0, DV != 0: vertical density diffusion (:261-264)
    EXT_REM = 4,     // lRem: conduction of TT - TTM plus the mean-profile term (:297-326)
    EXT_SHR = 8,     // lShr: FT carried over, heating/advection terms skipped (:293,477-538)
    EXT_HEAT = 16,   // iTCon == 2: embedded Gaussian heat loss (:443-461)
    EXT_PLUME = 32,  // TP != 0 with iTCon 0/1: Gaussian top-wall temperature (:750-776)
    EXT_ROCOPY = 64, // iD != 0 or lRem: RO ghost planes read as the planes two in (:230-239)
    EXT_RLAX = 128,  // RLAX > 0: FT += CORRECT(z) * RLAX * TFAC (:322-442)
};

// The flags whose branches act on interior points (fluxes() terms); the
// others (the plume walls, the RO ghost copy) act on the wall planes only,
// so a run with none of these keeps the default interior kernel.
constexpr int EXT_INTERIOR = EXT_DH | EXT_DV | EXT_REM | EXT_SHR | EXT_HEAT | EXT_RLAX;

// Whether branch X is on in a kernel compiled for branch set EXTM: never
// outside the set; always in a kernel compiled for X alone, or for lRem in
// the iD-with-lRem set, which are launched only for runs whose interior
// branches are exactly that set (so both paths need not be held in
// registers); otherwise as the run's flags say.
template <int EXTM, int X>
__host__ __device__ constexpr bool ext_on(int flags)
{
    return (EXTM & X) != 0 &&
           (EXTM == X || (X == EXT_REM && EXTM == (EXT_DH | EXT_DV | EXT_REM)) || (flags & X) != 0);
}

// Operands of the off-default branches (device pointers, host-formed scalars).
struct ExtArgs {
    int flags = 0;
    int iD = 0;
    int lRem = 0, lShr = 0;
    double cdh = 0.0, cdv = 0.0;          // DH*ORE, DV*ORE
    // horizontal means of the stage input, [0, nz): RO (with the ghost planes
    // replaced by the planes two in, 3dmhdsub.h:230-239), [nz, 2nz): TT
    const double* __restrict__ means = nullptr;
    const double* __restrict__ wwz2 = nullptr;    // iD == 2 weights per plane (:246-252)
    double* __restrict__ ft = nullptr;            // lShr: FT of the last fluxes() call (field layout)
    const double* __restrict__ heat_x = nullptr;  // iTCon 2: exp(CLN*(EXX-XP)^2/HH), per ghosted x
    const double* __restrict__ heat_y = nullptr;  //   ... per ghosted y
    const double* __restrict__ heat_z = nullptr;  //   ... per plane
    double heat_amp = 0.0;                        //   TP*OCV[*0.5f*(1+tanh(..))] of this step
    const double* __restrict__ plume_x = nullptr; // TP plume: exp(CLN*(EXX-XP)^2), per ghosted x
    const double* __restrict__ plume_y = nullptr; //   ... per ghosted y
    double tpr = 0.0;                             //   TPR of this step (3dmhdsub.h:752-757)
    // RLAX: CORRECT per plane [0, nz) and the factor RLAX * TFAC at [nz],
    // formed on the device before each fluxes() evaluation
    const double* __restrict__ relax = nullptr;
    double TU = 0.0, DZTU = 0.0, DZTB = 0.0;      // lRem wall temperatures (3dmhdphsic.h:590-592)
    double ORZD = 0.0;                            // 1.0f / double(nPz - 1) (3dmhdsub.h:768)
    double HH = 0.0, TIMI = 0.0;
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Order-preserving map of doubles onto uint64 (for atomic min/max).
__device__ __forceinline__ unsigned long long ord_key(double v)
{
    unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double ord_val(unsigned long long k)
{
    unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double d;
    __builtin_memcpy(&d, &u, sizeof(d));
    return d;
}
__host__ __device__ inline unsigned long long ord_key_host(double v)
{
    unsigned long long u;
    __builtin_memcpy(&u, &v, sizeof(u));
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// std::min / std::max of the reference (3dmhdutils.h:61-95): a NaN candidate
// never replaces the running value.
__host__ __device__ __forceinline__ double ref_min(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double ref_max(double a, double b) { return (a < b) ? b : a; }

}  // namespace mhd
