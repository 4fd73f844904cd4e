// Point-parallel kernels: wall planes of a stage, CFL reduction, BCON,
// stand-alone fluxes, diagnostics and small bookkeeping kernels.
//
// The wall kernel evaluates the one-sided stencils of the two z walls and
// then applies BCON (3dmhdsub.h:720-873) to the freshly updated wall values,
// reading the new values two planes inward that the interior stage kernel has
// already written.
#pragma once
#pragma clang fp contract(off)

#include "flux_point.hpp"
#include "mhd3d_types.hpp"

namespace mhd {

// State accessor that reads the register straight from global memory and
// forms 1/RO, u, v, w and P on the fly (3dmhdsub.h:267-276).
struct GAcc {
    const double* __restrict__ U;
    int64_t fs, plane;
    int pitch, nPx, nPy, x, y, z;
    __device__ __forceinline__ double ld(int f, int dx, int dy, int dz) const
    {
        const int xx = wrapi(x + dx, nPx), yy = wrapi(y + dy, nPy);
        return U[f * fs + (int64_t)(z + dz) * plane + (int64_t)yy * pitch + xx];
    }
    __device__ __forceinline__ double ru(int a, int b, int c) const { return ld(0, a, b, c); }
    __device__ __forceinline__ double rv(int a, int b, int c) const { return ld(1, a, b, c); }
    __device__ __forceinline__ double rw(int a, int b, int c) const { return ld(2, a, b, c); }
    __device__ __forceinline__ double ro(int a, int b, int c) const { return ld(3, a, b, c); }
    __device__ __forceinline__ double tt(int a, int b, int c) const { return ld(4, a, b, c); }
    __device__ __forceinline__ double bx(int a, int b, int c) const { return ld(5, a, b, c); }
    __device__ __forceinline__ double by(int a, int b, int c) const { return ld(6, a, b, c); }
    __device__ __forceinline__ double bz(int a, int b, int c) const { return ld(7, a, b, c); }
    __device__ __forceinline__ double ri(int a, int b, int c) const { return 1.0 / ro(a, b, c); }
    __device__ __forceinline__ double u(int a, int b, int c) const { return ru(a, b, c) * ri(a, b, c); }
    __device__ __forceinline__ double v(int a, int b, int c) const { return rv(a, b, c) * ri(a, b, c); }
    __device__ __forceinline__ double w(int a, int b, int c) const { return rw(a, b, c) * ri(a, b, c); }
    __device__ __forceinline__ double p(int a, int b, int c) const { return ro(a, b, c) * tt(a, b, c); }
};

struct CflPart;

struct StageArgs {
    Geom g;
    Phys c;
    Metric m;
    const double* __restrict__ Uin;
    double* __restrict__ Uout;
    double* __restrict__ Z;
    const double* __restrict__ dt;
    // stage 3 only: when non-null, the kernel also evaluates the CFL limits
    // of the new state (the TSTEP of the next step, 3dmhdsub.h:36-106) and
    // writes one partial record per block at cfl_parts[cfl_base + block]
    CflPart* cfl_parts;
    int cfl_base;
};

__device__ __forceinline__ PointMetric point_metric(const Metric& m, int xg, int yg, int zg)
{
    PointMetric pm;
    pm.dxx = m.dxxdx[xg];
    pm.d2xx = m.d2xxdx2[xg];
    pm.dyy = m.dyydy[yg];
    pm.d2yy = m.d2yydy2[yg];
    pm.dzz = m.dzzdz[zg];
    pm.d2zz = m.d2zzdz2[zg];
    pm.kap = m.RKAPA[zg];
    pm.dkap = m.DKAPA[zg];
    return pm;
}

__device__ __forceinline__ double fsel(const FluxOut& o, int f)
{
    switch (f) {
    case 0: return o.FU;
    case 1: return o.FV;
    case 2: return o.FW;
    case 3: return o.FR;
    case 4: return o.FT;
    case 5: return o.B1;
    case 6: return o.B2;
    default: return o.B3;
    }
}

// Per-point TSTEP quantities from the state values (3dmhdsub.h:41-88).
__device__ __forceinline__ void cfl_values(const Phys& c, const Metric& m, bool mag, double RU,
                                           double RV, double RW, double RO, double TT,
                                           double BX, double BY, double BZ, int xg, int yg,
                                           int z, double w[4], double& um)
{
    double UU = (RU * RU + RV * RV + RW * RW) / (RO * RO);
    um = c.OGAMMA * UU / TT;
    if (mag) {
        const double Q = c.GAMMA * TT + (BX * BX + BY * BY + BZ * BZ) * c.OBETA / RO;
        UU += Q + 2.0 * sqrt(UU * Q);
    } else {
        UU += c.GAMMA * TT + 2.0 * sqrt(c.GAMMA * UU * TT);
    }
    double DD = ref_min(m.DDY[yg], m.DDZ[z]);
    if (c.xdim) DD = ref_min(DD, m.DDX[xg]);
    w[0] = DD / sqrt(UU);
    w[1] = 0.5 * DD * DD * c.REPR * c.CV * RO / m.RKAPA[z];
    w[2] = 0.375 * DD * DD * c.RE * RO;
    w[3] = mag ? 0.5 * DD * DD * c.RM : 1e99;
}

__device__ __forceinline__ void cfl_point(const Geom& g, const Phys& c, const Metric& m,
                                          const double* __restrict__ U, bool mag, int x, int y,
                                          int z, double w[4], double& um)
{
    const int64_t i = (int64_t)z * g.plane + (int64_t)y * g.pitch + x;
    const int64_t fs = g.fs;
    const double bx = mag ? U[5 * fs + i] : 0.0;
    const double by = mag ? U[6 * fs + i] : 0.0;
    const double bz = mag ? U[7 * fs + i] : 0.0;
    cfl_values(c, m, mag, U[i], U[fs + i], U[2 * fs + i], U[3 * fs + i], U[4 * fs + i], bx, by,
               bz, x + 1, y + 1, z, w, um);
}

// Running CFL extrema of one thread; NaN candidates never replace the
// running value (std::min/std::max of 3dmhdutils.h:61-95).
struct CflAcc {
    double w[4] = {1e99, 1e99, 1e99, 1e99};
    double um = 0.0, bad = 0.0;
    __device__ __forceinline__ void add(const double wp[4], double up)
    {
        for (int k = 0; k < 4; ++k) w[k] = ref_min(w[k], wp[k]);
        um = ref_max(um, up);
        if (!(wp[0] - wp[0] == 0.0) || !(wp[1] - wp[1] == 0.0)) bad = 1.0;
    }
};

__device__ __forceinline__ double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = ref_min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = ref_max(v, __shfl_xor(v, o, 64));
    return v;
}

// One partial record per block: minima of WW1, WW2, WW3, VV and max of
// OGAMMA*UU/TT over the block's points, plus a non-finite flag.
struct CflPart {
    double w[4];
    double um;
    double bad;
};

// Block reduction of CflAcc into part[b] (wave shuffles, then LDS).
__device__ __forceinline__ void cfl_block_reduce(CflAcc& a, CflPart* part, int b)
{
    __shared__ double sh[6][16];
    for (int k = 0; k < 4; ++k) a.w[k] = wave_min(a.w[k]);
    a.um = wave_max(a.um);
    a.bad = wave_max(a.bad);
    const int tid = (threadIdx.z * blockDim.y + threadIdx.y) * blockDim.x + threadIdx.x;
    const int wv = tid >> 6, lane = tid & 63;
    if (lane == 0) {
        for (int k = 0; k < 4; ++k) sh[k][wv] = a.w[k];
        sh[4][wv] = a.um;
        sh[5][wv] = a.bad;
    }
    __syncthreads();
    if (tid == 0) {
        const int nw = (blockDim.x * blockDim.y * blockDim.z + 63) / 64;
        for (int j = 1; j < nw; ++j) {
            for (int k = 0; k < 4; ++k) a.w[k] = ref_min(a.w[k], sh[k][j]);
            a.um = ref_max(a.um, sh[4][j]);
            a.bad = ref_max(a.bad, sh[5][j]);
        }
        for (int k = 0; k < 4; ++k) part[b].w[k] = a.w[k];
        part[b].um = a.um;
        part[b].bad = a.bad;
    }
}

// ---------------------------------------------------------------------------
// Interior planes of one RK stage, one thread per point (baseline form).
// Stage k: F = fluxes(Uin); Uout = Zin + GAM_k*DT*F; Z = Uout + ZETA_k*DT*F
// (step_block_GAM / step_block_ZETA, 3dmhdsub.h:168-192; Zin = Uin for k = 1).
template <bool MAG, bool ROT, int STAGE>
__global__ void __launch_bounds__(256) stage_point_kernel(StageArgs a)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z + 2;   // planes 2 .. nPz-1
    if (x >= a.g.nPx || y >= a.g.nPy || z > a.g.nPz - 1) return;
    GAcc acc{a.Uin, a.g.fs, a.g.plane, a.g.pitch, a.g.nPx, a.g.nPy, x, y, z};
    const PointMetric pm = point_metric(a.m, x + 1, y + 1, z);
    FluxOut o;
    flux_point<MAG, ROT, 0, double>(acc, a.c, pm, o);
    const double DT = *a.dt;
    const double cg = a.c.GAM[STAGE - 1] * DT;
    const double cz = STAGE < 3 ? a.c.ZETA[STAGE - 1] * DT : 0.0;
    const int64_t i = (int64_t)z * a.g.plane + (int64_t)y * a.g.pitch + x;
    constexpr int NF = MAG ? NF_MAG : NF_HYDRO;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const double F = fsel(o, f);
        const double zin = STAGE == 1 ? a.Uin[f * a.g.fs + i] : a.Z[f * a.g.fs + i];
        const double un = zin + cg * F;
        a.Uout[f * a.g.fs + i] = un;
        if (STAGE < 3) a.Z[f * a.g.fs + i] = un + cz * F;
    }
}

// ---------------------------------------------------------------------------
// Wall planes of one RK stage: fluxes with the one-sided wall stencils, the
// GAM update, BCON on the updated wall values, then the ZETA register.
template <bool MAG, bool ROT, int STAGE>
__global__ void __launch_bounds__(256) stage_wall_kernel(StageArgs a)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const bool top = blockIdx.z == 0;
    // every thread reaches the block reduction below; threads outside the
    // domain only contribute the identity
    const bool active = x < a.g.nPx && y < a.g.nPy;
    constexpr int NF = MAG ? NF_MAG : NF_HYDRO;
    CflAcc cfl;
    if (active) {
        const int z = top ? 1 : a.g.nPz;
        const int zi1 = top ? z + 1 : z - 1;   // first plane inward
        const int zi2 = top ? z + 2 : z - 2;   // second plane inward
        GAcc acc{a.Uin, a.g.fs, a.g.plane, a.g.pitch, a.g.nPx, a.g.nPy, x, y, z};
        const PointMetric pm = point_metric(a.m, x + 1, y + 1, z);
        FluxOut o;
        if (top)
            flux_point<MAG, ROT, 1, double>(acc, a.c, pm, o);
        else
            flux_point<MAG, ROT, 2, double>(acc, a.c, pm, o);
        const double DT = *a.dt;
        const double cg = a.c.GAM[STAGE - 1] * DT;
        const double cz = STAGE < 3 ? a.c.ZETA[STAGE - 1] * DT : 0.0;
        const int64_t row = (int64_t)y * a.g.pitch + x;
        const int64_t i = (int64_t)z * a.g.plane + row;
        const int64_t i1 = (int64_t)zi1 * a.g.plane + row;
        const int64_t i2 = (int64_t)zi2 * a.g.plane + row;
        const int64_t fs = a.g.fs;
        double un[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const double zin = STAGE == 1 ? a.Uin[f * fs + i] : a.Z[f * fs + i];
            un[f] = zin + cg * fsel(o, f);
        }
        // BCON (3dmhdsub.h:727-872); the updated planes inward are in Uout
        const Phys& c = a.c;
        const double* Un = a.Uout;
        const double ro1 = Un[3 * fs + i1], ro2 = Un[3 * fs + i2];
        un[0] = un[3] * (c.C43 * Un[0 * fs + i1] / ro1 - c.C13 * Un[0 * fs + i2] / ro2);
        un[1] = un[3] * (c.C43 * Un[1 * fs + i1] / ro1 - c.C13 * Un[1 * fs + i2] / ro2);
        un[2] = 0.0;
        if (top) {
            if (c.iTCon == 0 || c.iTCon == 4) {
                un[4] = 1.0;
                if (c.iTCon == 4 && (x + 1) < (a.g.nPx + 2) / 2) un[4] = 1.0 - c.TP;
            } else if (c.iTCon == 1) {
                const double DZ = c.ORZF / a.m.dzzdz[z];
                un[4] = c.C43 * Un[4 * fs + i1] - c.C13 * Un[4 * fs + i2] - c.C23 * DZ * c.THETA;
            }
        } else {
            if (c.izCon == 0) {
                un[4] = c.TB;
            } else {
                const double DZ = c.ORZF / a.m.dzzdz[z];
                un[4] = c.C43 * Un[4 * fs + i1] - c.C13 * Un[4 * fs + i2] -
                        c.C23 * DZ * c.THETA / a.m.RKAPA[z];
            }
        }
        if (MAG && c.iBCon == 0) un[7] = 0.0;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            a.Uout[f * fs + i] = un[f];
            if (STAGE < 3) a.Z[f * fs + i] = un[f] + cz * fsel(o, f);
        }
        if (STAGE == 3 && a.cfl_parts) {
            double wp[4], up;
            cfl_values(c, a.m, MAG, un[0], un[1], un[2], un[3], un[4], MAG ? un[5] : 0.0,
                       MAG ? un[6] : 0.0, MAG ? un[7] : 0.0, x + 1, y + 1, z, wp, up);
            cfl.add(wp, up);
        }
    }
    if (STAGE == 3 && a.cfl_parts) {
        const int b = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        cfl_block_reduce(cfl, a.cfl_parts, a.cfl_base + b);
    }
}

// ---------------------------------------------------------------------------
// Stand-alone fluxes() into F (all interior planes incl. walls).
template <bool MAG, bool ROT>
__global__ void __launch_bounds__(256) flux_only_kernel(StageArgs a, double* __restrict__ F)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z + 1;   // planes 1 .. nPz
    if (x >= a.g.nPx || y >= a.g.nPy || z > a.g.nPz) return;
    GAcc acc{a.Uin, a.g.fs, a.g.plane, a.g.pitch, a.g.nPx, a.g.nPy, x, y, z};
    const PointMetric pm = point_metric(a.m, x + 1, y + 1, z);
    FluxOut o;
    if (z == 1)
        flux_point<MAG, ROT, 1, double>(acc, a.c, pm, o);
    else if (z == a.g.nPz)
        flux_point<MAG, ROT, 2, double>(acc, a.c, pm, o);
    else
        flux_point<MAG, ROT, 0, double>(acc, a.c, pm, o);
    const int64_t i = (int64_t)z * a.g.plane + (int64_t)y * a.g.pitch + x;
    constexpr int NF = MAG ? NF_MAG : NF_HYDRO;
#pragma unroll
    for (int f = 0; f < NF; ++f) F[f * a.g.fs + i] = fsel(o, f);
}

// ---------------------------------------------------------------------------
// BCON() on a register in place (3dmhdsub.h:720-873).
template <bool MAG>
__global__ void bcon_kernel(Geom g, Phys c, Metric m, double* __restrict__ U)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const bool top = blockIdx.z == 0;
    if (x >= g.nPx || y >= g.nPy) return;
    const int z = top ? 1 : g.nPz;
    const int64_t row = (int64_t)y * g.pitch + x;
    const int64_t i = (int64_t)z * g.plane + row;
    const int64_t i1 = (int64_t)(top ? z + 1 : z - 1) * g.plane + row;
    const int64_t i2 = (int64_t)(top ? z + 2 : z - 2) * g.plane + row;
    const int64_t fs = g.fs;
    const double ro = U[3 * fs + i], ro1 = U[3 * fs + i1], ro2 = U[3 * fs + i2];
    U[0 * fs + i] = ro * (c.C43 * U[0 * fs + i1] / ro1 - c.C13 * U[0 * fs + i2] / ro2);
    U[1 * fs + i] = ro * (c.C43 * U[1 * fs + i1] / ro1 - c.C13 * U[1 * fs + i2] / ro2);
    U[2 * fs + i] = 0.0;
    if (top) {
        if (c.iTCon == 0 || c.iTCon == 4) {
            double t = 1.0;
            if (c.iTCon == 4 && (x + 1) < (g.nPx + 2) / 2) t = 1.0 - c.TP;
            U[4 * fs + i] = t;
        } else if (c.iTCon == 1) {
            const double DZ = c.ORZF / m.dzzdz[z];
            U[4 * fs + i] = c.C43 * U[4 * fs + i1] - c.C13 * U[4 * fs + i2] - c.C23 * DZ * c.THETA;
        }
    } else {
        if (c.izCon == 0) {
            U[4 * fs + i] = c.TB;
        } else {
            const double DZ = c.ORZF / m.dzzdz[z];
            U[4 * fs + i] = c.C43 * U[4 * fs + i1] - c.C13 * U[4 * fs + i2] -
                            c.C23 * DZ * c.THETA / m.RKAPA[z];
        }
    }
    if (MAG && c.iBCon == 0) U[7 * fs + i] = 0.0;
}

// ---------------------------------------------------------------------------
// TSTEP (3dmhdsub.h:36-106): per-point CFL limits, then a wave + block
// reduction into one partial record per block; cfl_finalize_kernel reduces
// the partials.  Min and max are exact, so the result does not depend on the
// reduction order.

template <bool MAG>
__global__ void __launch_bounds__(256) cfl_kernel(Geom g, Phys c, Metric m,
                                                  const double* __restrict__ U,
                                                  CflPart* __restrict__ part)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    CflAcc acc;
    if (x < g.nPx && y < g.nPy) {
        for (int z = 1 + blockIdx.z; z <= g.nPz; z += gridDim.z) {
            double wp[4], up;
            cfl_point(g, c, m, U, MAG, x, y, z, wp, up);
            acc.add(wp, up);
        }
    }
    const int b = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    cfl_block_reduce(acc, part, b);
}

// Device scalars of one context.
struct DevScalars {
    double dt;          // DT of the current step
    double timc;        // TIMC (3dmhdsub.h:151)
    double dt_last;     // DT of the last completed step
    double nit;         // nIt
    double umach;
    double wmin[4];
    unsigned int nonfinite;
    unsigned int pad;
};

// Reduces the per-block partials in a fixed order (one block of 256 threads:
// strided per-thread reduction, then the block reduction) and forms
// DT = SFF * min(...) as 3dmhdsub.h:101-105.
__global__ void __launch_bounds__(256) cfl_finalize_kernel(const CflPart* __restrict__ part,
                                                           int nparts, DevScalars* s, Phys c,
                                                           int mag)
{
    __shared__ double sh[6][256];
    double w[4] = {1e99, 1e99, 1e99, 1e99};
    double um = 0.0, bad = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        for (int k = 0; k < 4; ++k) w[k] = ref_min(w[k], part[b].w[k]);
        um = ref_max(um, part[b].um);
        bad = ref_max(bad, part[b].bad);
    }
    for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = w[k];
    sh[4][threadIdx.x] = um;
    sh[5][threadIdx.x] = bad;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            for (int k = 0; k < 4; ++k)
                sh[k][threadIdx.x] = ref_min(sh[k][threadIdx.x], sh[k][threadIdx.x + st]);
            sh[4][threadIdx.x] = ref_max(sh[4][threadIdx.x], sh[4][threadIdx.x + st]);
            sh[5][threadIdx.x] = ref_max(sh[5][threadIdx.x], sh[5][threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; ++k) w[k] = sh[k][0];
        s->umach = sqrt(sh[4][0]);
        for (int k = 0; k < 4; ++k) s->wmin[k] = w[k];
        if (sh[5][0] != 0.0) s->nonfinite = 1u;
        if (mag)
            s->dt = c.SFF * ref_min(ref_min(w[0], w[1]), ref_min(w[2], w[3]));
        else
            s->dt = c.SFF * ref_min(ref_min(w[0], w[1]), w[2]);
    }
}

__global__ void set_dt_kernel(DevScalars* s, double dt) { s->dt = dt; }

__global__ void clock_kernel(DevScalars* s)
{
    s->timc += s->dt;
    s->nit += 1.0;
    s->dt_last = s->dt;
}

// Non-finite scan of a register (interior incl. walls): one thread per
// (x, y) column, looping over z.
__global__ void __launch_bounds__(256) nonfinite_kernel(Geom g, int nf,
                                                        const double* __restrict__ U,
                                                        DevScalars* s)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= g.nPx || y >= g.nPy) return;
    bool bad = false;
    for (int z = 1 + blockIdx.z; z <= g.nPz; z += gridDim.z) {
        const int64_t i = (int64_t)z * g.plane + (int64_t)y * g.pitch + x;
        for (int f = 0; f < nf; ++f) {
            const double v = U[f * g.fs + i];
            if (!(v - v == 0.0)) bad = true;
        }
    }
    if (bad) atomicOr(&s->nonfinite, 1u);
}

__global__ void clear_flag_kernel(DevScalars* s) { s->nonfinite = 0u; }

}  // namespace mhd

namespace mhd {

// Global FP64 diagnostics: total mass, total energy, max |div B|.
// Deterministic: fixed per-block partials, summed in block order.
__global__ void __launch_bounds__(256) diag_kernel(Geom g, Phys c, Metric m, int nf,
                                                   const double* __restrict__ U,
                                                   double* __restrict__ part)
{
    __shared__ double sh[3][256];
    double ms = 0.0, en = 0.0, db = 0.0;
    const int64_t npts = (int64_t)g.nPz * g.nPy * g.nPx;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < npts;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(t % g.nPx);
        const int y = (int)((t / g.nPx) % g.nPy);
        const int z = (int)(t / ((int64_t)g.nPx * g.nPy)) + 1;
        const int64_t i = (int64_t)z * g.plane + (int64_t)y * g.pitch + x;
        const double RU = U[i], RV = U[g.fs + i], RW = U[2 * g.fs + i];
        const double RO = U[3 * g.fs + i], TT = U[4 * g.fs + i];
        ms += RO;
        double e = 0.5 * (RU * RU + RV * RV + RW * RW) / RO + c.CV * RO * TT;
        if (nf == NF_MAG) {
            const double BX = U[5 * g.fs + i], BY = U[6 * g.fs + i], BZ = U[7 * g.fs + i];
            e += 0.5 * c.OBETA * (BX * BX + BY * BY + BZ * BZ);
            if (z > 1 && z < g.nPz) {
                const int xp = wrapi(x + 1, g.nPx), xm = wrapi(x - 1, g.nPx);
                const int yp = wrapi(y + 1, g.nPy), ym = wrapi(y - 1, g.nPy);
                const int64_t r = (int64_t)z * g.plane;
                const double dbx = (U[5 * g.fs + r + (int64_t)y * g.pitch + xp] -
                                    U[5 * g.fs + r + (int64_t)y * g.pitch + xm]) * m.dxxdx[x + 1] * c.hx;
                const double dby = (U[6 * g.fs + r + (int64_t)yp * g.pitch + x] -
                                    U[6 * g.fs + r + (int64_t)ym * g.pitch + x]) * m.dyydy[y + 1] * c.hy;
                const double dbz = (U[7 * g.fs + i + g.plane] - U[7 * g.fs + i - g.plane]) *
                                   m.dzzdz[z] * c.hz;
                db = fmax(db, fabs(dbx + dby + dbz));
            }
        }
        en += e;
    }
    sh[0][threadIdx.x] = ms;
    sh[1][threadIdx.x] = en;
    sh[2][threadIdx.x] = db;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
            sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = sh[0][0];
        part[3 * blockIdx.x + 1] = sh[1][0];
        part[3 * blockIdx.x + 2] = sh[2][0];
    }
}

__global__ void diag_finalize_kernel(const double* part, int n, double* out)
{
    if (threadIdx.x != 0) return;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int k = 0; k < n; ++k) {
        a += part[3 * k];
        b += part[3 * k + 1];
        c = fmax(c, part[3 * k + 2]);
    }
    out[0] = a;
    out[1] = b;
    out[2] = c;
}

}  // namespace mhd
