// z-streaming RK stage kernel for the planes between the walls.
//
// One workgroup = a 32x8 tile of (x, y) columns and a chunk of z planes; one
// thread per column.  Each plane of the state is read from HBM once per tile
// (plus a 1-point halo ring): the thread that loads a point forms 1/RO, u, v,
// w and P = RO*TT there (the reference's extended-range sweeps
// 3dmhdsub.h:266-276) and stores them in LDS:
//   * u, v, w (and BX, BY, BZ) for three planes (z-1, z, z+1) -- the mixed
//     derivatives D1x(D1z w), D1y(D1z v), ... need u, v, w at (x+-1, z+-1),
//     (y+-1, z+-1);
//   * RU, RV, RW, TT, P for the centre plane only, double buffered so the
//     next plane can be written while the current one is read.
// The z neighbours of RU, RV, RW, TT, P at a thread's own column stay in
// registers.  The flux is
// the shared flux_point() (same operation order as the reference), followed by
// the fused step_block_GAM / step_block_ZETA update (3dmhdsub.h:168-192).
// One barrier per plane (four-slot ring, three centre-plane buffers) or two
// (three-slot ring, two buffers); global loads of the next plane are issued
// before the flux evaluation so they overlap it.
#pragma once
#pragma clang fp contract(off)

#include "kernels_basic.hpp"

namespace mhd {

constexpr int STX = 32, STY = 8;

// Workgroup barrier ordering LDS only.  Threads of a workgroup exchange data
// through LDS alone (every global location is read or written by one
// thread), so the barrier need not wait for outstanding global loads and
// stores, which __syncthreads() may do.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
constexpr int SHW = STX + 2, SHH = STY + 2, SHP = SHW * SHH;   // 34 x 10 = 340 points
constexpr int SNH = 2 * SHW + 2 * STY;                          // 84 halo points

// Raw values of one point of one plane.
template <bool MAG>
struct RawPt {
    double ru = 0.0, rv = 0.0, rw = 0.0, ro = 1.0, tt = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
};

template <bool MAG>
__device__ __forceinline__ void load_raw(RawPt<MAG>& r, const double* __restrict__ U,
                                         int64_t fs, int64_t off)
{
    r.ru = U[off];
    r.rv = U[fs + off];
    r.rw = U[2 * fs + off];
    r.ro = U[3 * fs + off];
    r.tt = U[4 * fs + off];
    if (MAG) {
        r.bx = U[5 * fs + off];
        r.by = U[6 * fs + off];
        r.bz = U[7 * fs + off];
    }
}

// LDS-backed accessor for an interior plane.  Star fields are
// [RU, RV, RW, TT, P, BX, BY, BZ]; ws[f][k] is the own-column window
// (k = 0, 1, 2 for z-1, z, z+1).
template <bool MAG>
struct SAcc {
    static constexpr int NW = 5;               // RU RV RW TT P
    const double* st;                          // star plane z, at this thread's centre
    const double* um;                          // u/v/w [BX BY BZ] ring, plane z-1, at centre
    const double* u0;                          // plane z
    const double* up;                          // plane z+1
    double ws[NW][3];
    double ro0, ri0;

    __device__ __forceinline__ double star(int f, int dx, int dy, int dz) const
    {
        if (dx == 0 && dy == 0) return ws[f][dz + 1];
        return st[f * SHP + dy * SHW + dx];
    }
    // u, v, w (and BX, BY, BZ) come from the three-plane LDS ring
    __device__ __forceinline__ double ring(int f, int dx, int dy, int dz) const
    {
        const double* b = dz < 0 ? um : (dz > 0 ? up : u0);
        return b[f * SHP + dy * SHW + dx];
    }
    __device__ __forceinline__ double ru(int a, int b, int c) const { return star(0, a, b, c); }
    __device__ __forceinline__ double rv(int a, int b, int c) const { return star(1, a, b, c); }
    __device__ __forceinline__ double rw(int a, int b, int c) const { return star(2, a, b, c); }
    __device__ __forceinline__ double tt(int a, int b, int c) const { return star(3, a, b, c); }
    __device__ __forceinline__ double p(int a, int b, int c) const { return star(4, a, b, c); }
    __device__ __forceinline__ double u(int a, int b, int c) const { return ring(0, a, b, c); }
    __device__ __forceinline__ double v(int a, int b, int c) const { return ring(1, a, b, c); }
    __device__ __forceinline__ double w(int a, int b, int c) const { return ring(2, a, b, c); }
    __device__ __forceinline__ double bx(int a, int b, int c) const { return ring(3, a, b, c); }
    __device__ __forceinline__ double by(int a, int b, int c) const { return ring(4, a, b, c); }
    __device__ __forceinline__ double bz(int a, int b, int c) const { return ring(5, a, b, c); }
    __device__ __forceinline__ double ro(int, int, int) const { return ro0; }
    __device__ __forceinline__ double ri(int, int, int) const { return ri0; }
};

// WAVES: minimum waves per SIMD the register allocation must allow (2 keeps
// 256 VGPRs per wave; 1 lets the MHD variant use AGPRs instead of spilling).
template <bool MAG, bool ROT, int STAGE, int WAVES, bool ONE_BARRIER>
__global__ void __launch_bounds__(256, WAVES) stage_stream_kernel(StageArgs a, int zchunk, int ntx,
                                                                  int nty, int xcd_remap)
{
    constexpr int NR = MAG ? 6 : 3;   // ring fields: u v w [BX BY BZ]
    constexpr int NS = 5;             // centre-plane fields: RU RV RW TT P
    constexpr int NF = MAG ? NF_MAG : NF_HYDRO;
    // ring depths: NRING planes of u, v, w (BX, BY, BZ) and NSTAR centre
    // planes, deep enough that one barrier per plane suffices (a thread can be
    // at most one plane ahead of another, and then it writes slots nobody
    // still reads)
    constexpr int NRING = ONE_BARRIER ? 4 : 3;
    constexpr int NSTAR = ONE_BARRIER ? 3 : 2;
    __shared__ double s_uvw[NRING][NR][SHP];
    __shared__ double s_st[NSTAR][NS][SHP];

    const Geom& g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int t = ty * STX + tx;

    // Block -> tile: with xcd_remap, consecutive tiles go to the same XCD
    // (blocks are dealt to the 8 XCDs round-robin) so halo re-reads hit its L2.
    const int nb = gridDim.x;
    const int bid = blockIdx.x;
    const int logical = xcd_remap ? (bid % 8) * (nb / 8) + bid / 8 : bid;
    const int tile_x = logical % ntx;
    const int rest = logical / ntx;
    const int tile_y = rest % nty;
    const int chunk = rest / nty;
    const int x0 = tile_x * STX, y0 = tile_y * STY;
    const int zb = 2 + chunk * zchunk;
    const int ze = min(zb + zchunk, g.nPz);   // planes zb .. ze-1 (interior 2 .. nPz-1)
    if (zb >= ze) {
        if (STAGE == 3 && a.cfl_parts && threadIdx.x == 0 && threadIdx.y == 0) {
            CflPart& q = a.cfl_parts[a.cfl_base + blockIdx.x];
            for (int k = 0; k < 4; ++k) q.w[k] = 1e99;
            q.um = 0.0;
            q.bad = 0.0;
        }
        return;
    }

    // own column and (for t < SNH) one halo point, wrapped periodically
    const int gx = (x0 + tx) % g.nPx, gy = (y0 + ty) % g.nPy;
    const bool valid = (x0 + tx < g.nPx) && (y0 + ty < g.nPy);
    const int64_t go = (int64_t)gy * g.pitch + gx;
    const int pc = (ty + 1) * SHW + (tx + 1);
    int ph = 0;
    int64_t gh = 0;
    const bool has_halo = t < SNH;
    if (has_halo) {
        int lx, ly;
        if (t < SHW) { lx = t; ly = 0; }
        else if (t < 2 * SHW) { lx = t - SHW; ly = SHH - 1; }
        else if (t < 2 * SHW + STY) { lx = 0; ly = t - 2 * SHW + 1; }
        else { lx = SHW - 1; ly = t - 2 * SHW - STY + 1; }
        const int hx = (x0 - 1 + lx + g.nPx) % g.nPx;
        const int hy = (y0 - 1 + ly + g.nPy) % g.nPy;
        gh = (int64_t)hy * g.pitch + hx;
        ph = ly * SHW + lx;
    }

    const double* __restrict__ U = a.Uin;
    const int64_t fs = g.fs;
    SAcc<MAG> acc;
    double row[3] = {0.0, 0.0, 0.0};
    double rinv[3] = {0.0, 0.0, 0.0};
    RawPt<MAG> own, halo, own_n, halo_n;
    load_raw<MAG>(own, U, fs, (int64_t)(zb - 1) * g.plane + go);
    if (has_halo) load_raw<MAG>(halo, U, fs, (int64_t)(zb - 1) * g.plane + gh);

    const double DT = *a.dt;
    const double cg = a.c.GAM[STAGE - 1] * DT;
    const double cz = STAGE < 3 ? a.c.ZETA[STAGE - 1] * DT : 0.0;
    // Z of the next plane to update, loaded one iteration ahead (stages 2, 3)
    double zpre[NF], zcur[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) zpre[f] = zcur[f] = 0.0;
    if (STAGE > 1) {
#pragma unroll
        for (int f = 0; f < NF; ++f) zpre[f] = a.Z[f * fs + (int64_t)zb * g.plane + go];
    }

    CflAcc cfl;   // stage 3: TSTEP of the next step on the new state
    for (int L = zb - 1; L <= ze; ++L) {
        // -- derive plane L and publish it
        {
            const double ri = 1.0 / own.ro;
            const double u = own.ru * ri, v = own.rv * ri, w = own.rw * ri, P = own.ro * own.tt;
            double* su = &s_uvw[L % NRING][0][0];
            double* ss = &s_st[L % NSTAR][0][0];
            su[0 * SHP + pc] = u;
            su[1 * SHP + pc] = v;
            su[2 * SHP + pc] = w;
            ss[0 * SHP + pc] = own.ru;
            ss[1 * SHP + pc] = own.rv;
            ss[2 * SHP + pc] = own.rw;
            ss[3 * SHP + pc] = own.tt;
            ss[4 * SHP + pc] = P;
            if (MAG) {
                su[3 * SHP + pc] = own.bx;
                su[4 * SHP + pc] = own.by;
                su[5 * SHP + pc] = own.bz;
            }
            const double nw[5] = {own.ru, own.rv, own.rw, own.tt, P};
#pragma unroll
            for (int k = 0; k < SAcc<MAG>::NW; ++k) {
                acc.ws[k][0] = acc.ws[k][1];
                acc.ws[k][1] = acc.ws[k][2];
                acc.ws[k][2] = nw[k];
            }
            row[0] = row[1]; row[1] = row[2]; row[2] = own.ro;
            rinv[0] = rinv[1]; rinv[1] = rinv[2]; rinv[2] = ri;
            if (has_halo) {
                const double hri = 1.0 / halo.ro;
                su[0 * SHP + ph] = halo.ru * hri;
                su[1 * SHP + ph] = halo.rv * hri;
                su[2 * SHP + ph] = halo.rw * hri;
                ss[0 * SHP + ph] = halo.ru;
                ss[1 * SHP + ph] = halo.rv;
                ss[2 * SHP + ph] = halo.rw;
                ss[3 * SHP + ph] = halo.tt;
                ss[4 * SHP + ph] = halo.ro * halo.tt;
                if (MAG) {
                    su[3 * SHP + ph] = halo.bx;
                    su[4 * SHP + ph] = halo.by;
                    su[5 * SHP + ph] = halo.bz;
                }
            }
        }
        lds_barrier();
        // -- prefetch plane L+1 (before the flux evaluation, so the loads overlap it)
        if (L < ze) {
            load_raw<MAG>(own_n, U, fs, (int64_t)(L + 1) * g.plane + go);
            if (has_halo) load_raw<MAG>(halo_n, U, fs, (int64_t)(L + 1) * g.plane + gh);
        }
        const int z = L - 1;
        if (STAGE > 1 && z >= zb) {
            // this iteration updates plane z; fetch Z of plane z+1 for the next one
#pragma unroll
            for (int f = 0; f < NF; ++f) zcur[f] = zpre[f];
            if (z + 1 < ze) {
#pragma unroll
                for (int f = 0; f < NF; ++f) zpre[f] = a.Z[f * fs + (int64_t)(z + 1) * g.plane + go];
            }
        }
        if (z >= zb) {
            acc.st = &s_st[z % NSTAR][0][pc];
            acc.um = &s_uvw[(z - 1) % NRING][0][pc];
            acc.u0 = &s_uvw[z % NRING][0][pc];
            acc.up = &s_uvw[(z + 1) % NRING][0][pc];
            acc.ro0 = row[1];
            acc.ri0 = rinv[1];
            const PointMetric pm = point_metric(a.m, gx + 1, gy + 1, z);
            FluxOut o;
            flux_point<MAG, ROT, 0, double>(acc, a.c, pm, o);
            if (valid) {
                const int64_t i = (int64_t)z * g.plane + go;
                double uc[8] = {acc.ws[0][1], acc.ws[1][1], acc.ws[2][1], row[1],
                                acc.ws[3][1], 0.0, 0.0, 0.0};
                if constexpr (MAG) {
                    uc[5] = acc.u0[3 * SHP];
                    uc[6] = acc.u0[4 * SHP];
                    uc[7] = acc.u0[5 * SHP];
                }
                double un[NF];
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    const double F = fsel(o, f);
                    const double zin = STAGE == 1 ? uc[f] : zcur[f];
                    un[f] = zin + cg * F;
                    a.Uout[f * fs + i] = un[f];
                    if (STAGE < 3) a.Z[f * fs + i] = un[f] + cz * F;
                }
                if (STAGE == 3 && a.cfl_parts) {
                    double wp[4], up;
                    if constexpr (MAG)
                        cfl_values(a.c, a.m, true, un[0], un[1], un[2], un[3], un[4], un[5],
                                   un[6], un[7], gx + 1, gy + 1, z, wp, up);
                    else
                        cfl_values(a.c, a.m, false, un[0], un[1], un[2], un[3], un[4], 0.0,
                                   0.0, 0.0, gx + 1, gy + 1, z, wp, up);
                    cfl.add(wp, up);
                }
            }
        }
        if (!ONE_BARRIER) lds_barrier();
        own = own_n;
        if (has_halo) halo = halo_n;
    }
    if (STAGE == 3 && a.cfl_parts) cfl_block_reduce(cfl, a.cfl_parts, a.cfl_base + blockIdx.x);
}

struct StreamPlan {
    int zchunk, nchunks, ntx, nty, nblocks, xcd_remap;
};

inline StreamPlan stream_plan(const Geom& g, int target_blocks = 1024)
{
    StreamPlan p;
    p.ntx = (g.nPx + STX - 1) / STX;
    p.nty = (g.nPy + STY - 1) / STY;
    const int nint = g.nPz - 2;
    const int tiles = p.ntx * p.nty;
    int nch = (target_blocks + tiles - 1) / tiles;
    const int maxch = nint / 8 > 0 ? nint / 8 : 1;
    if (nch > maxch) nch = maxch;
    if (nch < 1) nch = 1;
    p.zchunk = (nint + nch - 1) / nch;
    p.nchunks = (nint + p.zchunk - 1) / p.zchunk;
    p.nblocks = tiles * p.nchunks;
    p.xcd_remap = (p.nblocks % 8 == 0) ? 1 : 0;
    return p;
}

template <bool MAG, bool ROT, int STAGE>
static void launch_stream_stage(const StageArgs& a, hipStream_t s, int waves, bool one_barrier)
{
    const StreamPlan p = stream_plan(a.g);
    dim3 blk(STX, STY, 1);
    if (waves == 1 && one_barrier)
        stage_stream_kernel<MAG, ROT, STAGE, 1, true><<<p.nblocks, blk, 0, s>>>(
            a, p.zchunk, p.ntx, p.nty, p.xcd_remap);
    else if (waves == 1)
        stage_stream_kernel<MAG, ROT, STAGE, 1, false><<<p.nblocks, blk, 0, s>>>(
            a, p.zchunk, p.ntx, p.nty, p.xcd_remap);
    else if (one_barrier)
        stage_stream_kernel<MAG, ROT, STAGE, 2, true><<<p.nblocks, blk, 0, s>>>(
            a, p.zchunk, p.ntx, p.nty, p.xcd_remap);
    else
        stage_stream_kernel<MAG, ROT, STAGE, 2, false><<<p.nblocks, blk, 0, s>>>(
            a, p.zchunk, p.ntx, p.nty, p.xcd_remap);
}

}  // namespace mhd
