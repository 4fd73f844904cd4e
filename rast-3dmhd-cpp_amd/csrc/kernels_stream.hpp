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
// FP64 arithmetic follows the reference's operation order, so nothing is
// contracted into FMA; the translation unit of the FP32 (mixed-form) kernels
// alone allows contraction within an expression.
#ifdef MHD3D_FP32_FORM_TU
#pragma clang fp contract(on)
#else
#pragma clang fp contract(off)
#endif

#include <cstddef>
#include <type_traits>

#include "kernels_basic.hpp"

namespace mhd {

constexpr int STX = 32, STY = 8;

// Store of a stage result nothing in the same launch reads again, marked
// nontemporal so the written planes do not displace from L2 the lines the
// neighbouring tiles still read as their halos (MI355X, FP64 MHD
// 512x512x256: stage kernel 3.65 -> 3.48 ms, HBM reads -3 to -10%).
template <typename V>
__device__ __forceinline__ void stream_store(V* p, V v)
{
    __builtin_nontemporal_store(v, p);
}

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

// Values of one point of one plane as stored in the stencil input (type TI).
template <typename TI, bool MAG>
struct RawPt {
    TI ru = TI(0), rv = TI(0), rw = TI(0), ro = TI(1), tt = TI(0), bx = TI(0), by = TI(0), bz = TI(0);
};

template <typename TI, bool MAG>
__device__ __forceinline__ void load_raw(RawPt<TI, MAG>& r, const TI* __restrict__ U, int64_t fs,
                                         int64_t off)
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

// Stencil values of one point in the type T: RU RV RW, the full density, the
// TT and P stencil values (deviations from the background in the mixed
// form), the density deviation and BX BY BZ.
template <typename T>
struct PrepPt {
    T ru, rv, rw, ro, tt, p, rdev, bx, by, bz;
};

// FP64 form: the stored values.  Mixed form: the stored FP32 deviations of
// RO, TT, BX, BY, BZ (or, when the input is the FP64 state, its deviations
// from the FP64 background rounded to FP32), with the full density RO0 + r
// and P - P0 = (RO0*t + r*TT0) + r*t formed from the FP32 background of
// plane z.  FP64 values from an FP32 copy (the mixed form with the
// off-default branches): the deviations plus the FP64 background.
template <typename T, bool MAG, typename TI>
__device__ __forceinline__ PrepPt<T> prep_point(const RawPt<TI, MAG>& raw, const StageArgs& a, int z)
{
    PrepPt<T> q;
    q.ru = T(raw.ru);
    q.rv = T(raw.rv);
    q.rw = T(raw.rw);
    if constexpr (!std::is_same<T, double>::value && std::is_same<TI, double>::value && MAG) {
        const int nz = a.g.nz;
        q.bx = T(raw.bx - uniform_load(a.bg64 + (BG64_B0 + 0) * nz + z));
        q.by = T(raw.by - uniform_load(a.bg64 + (BG64_B0 + 1) * nz + z));
        q.bz = T(raw.bz - uniform_load(a.bg64 + (BG64_B0 + 2) * nz + z));
    } else if constexpr (std::is_same<T, double>::value && std::is_same<TI, float>::value && MAG) {
        const int nz = a.g.nz;
        q.bx = double(raw.bx) + uniform_load(a.bg64 + (BG64_B0 + 0) * nz + z);
        q.by = double(raw.by) + uniform_load(a.bg64 + (BG64_B0 + 1) * nz + z);
        q.bz = double(raw.bz) + uniform_load(a.bg64 + (BG64_B0 + 2) * nz + z);
    } else {
        q.bx = T(raw.bx);
        q.by = T(raw.by);
        q.bz = T(raw.bz);
    }
    if constexpr (std::is_same<T, double>::value && std::is_same<TI, float>::value) {
        // FP64 fluxes from the mixed form's FP32 copy: the FP64 background
        // of plane z added back to the stored deviations
        const int nz = a.g.nz;
        q.ro = double(raw.ro) + uniform_load(a.bg64 + BG64_R0 * nz + z);
        q.tt = double(raw.tt) + uniform_load(a.bg64 + BG64_T0 * nz + z);
        q.p = q.ro * q.tt;
        q.rdev = T(0);
    } else if constexpr (std::is_same<T, double>::value) {
        q.ro = raw.ro;
        q.tt = raw.tt;
        q.p = raw.ro * raw.tt;
        q.rdev = T(0);
    } else {
        const int nz = a.g.nz;
        const T R0 = uniform_load(a.bgf + BG_R0 * nz + z), T0 = uniform_load(a.bgf + BG_T0 * nz + z);
        T r, t;
        if constexpr (std::is_same<TI, double>::value) {
            r = T(raw.ro - uniform_load(a.bg64 + BG64_R0 * nz + z));
            t = T(raw.tt - uniform_load(a.bg64 + BG64_T0 * nz + z));
        } else {
            r = raw.ro;
            t = raw.tt;
        }
        q.ro = R0 + r;
        q.tt = t;
        q.p = (R0 * t + r * T0) + r * t;
        q.rdev = r;
    }
    return q;
}

// LDS-backed accessor for an interior plane.  Centre-plane fields are
// [RU, RV, RW, TT, P] and, with RO_N, RO; ws[f][k] is the own-column window
// (k = 0, 1, 2 for z-1, z, z+1).  RO_N: ro() at the neighbours too (the
// centre plane's RO from LDS, the own column's z-1 / z+1 from rz).
template <typename T, bool MAG, bool RO_N = false>
struct SAcc {
    static constexpr bool kNeighbourRO = RO_N;
    static constexpr int NW = 5;               // RU RV RW TT P
    const T* st;                               // centre plane z, at this thread's centre
    const T* um;                               // u/v/w [BX BY BZ] ring, plane z-1, at centre
    const T* u0;                               // plane z
    const T* up;                               // plane z+1
    T ws[NW][3];
    T ro0, ri0;
    T rz[3];

    __device__ __forceinline__ T star(int f, int dx, int dy, int dz) const
    {
        if (dx == 0 && dy == 0) return ws[f][dz + 1];
        return st[f * SHP + dy * SHW + dx];
    }
    // u, v, w (and BX, BY, BZ) come from the ring
    __device__ __forceinline__ T ring(int f, int dx, int dy, int dz) const
    {
        const T* b = dz < 0 ? um : (dz > 0 ? up : u0);
        return b[f * SHP + dy * SHW + dx];
    }
    __device__ __forceinline__ T ru(int a, int b, int c) const { return star(0, a, b, c); }
    __device__ __forceinline__ T rv(int a, int b, int c) const { return star(1, a, b, c); }
    __device__ __forceinline__ T rw(int a, int b, int c) const { return star(2, a, b, c); }
    __device__ __forceinline__ T tt(int a, int b, int c) const { return star(3, a, b, c); }
    __device__ __forceinline__ T p(int a, int b, int c) const { return star(4, a, b, c); }
    __device__ __forceinline__ T u(int a, int b, int c) const { return ring(0, a, b, c); }
    __device__ __forceinline__ T v(int a, int b, int c) const { return ring(1, a, b, c); }
    __device__ __forceinline__ T w(int a, int b, int c) const { return ring(2, a, b, c); }
    __device__ __forceinline__ T bx(int a, int b, int c) const { return ring(3, a, b, c); }
    __device__ __forceinline__ T by(int a, int b, int c) const { return ring(4, a, b, c); }
    __device__ __forceinline__ T bz(int a, int b, int c) const { return ring(5, a, b, c); }
    __device__ __forceinline__ T ro(int dx, int dy, int dz) const
    {
        if constexpr (RO_N) {
            if (dz != 0) return rz[dz + 1];
            if (dx != 0 || dy != 0) return st[5 * SHP + dy * SHW + dx];
        }
        return ro0;
    }
    __device__ __forceinline__ T ri(int, int, int) const { return ri0; }
    // deviation form (mixed): background terms of the centre plane
    T hb_ = T(0), dzt0_ = T(0), d2zt0_ = T(0), t0_ = T(0), rdev_ = T(0);
    T b0_[3] = {T(0), T(0), T(0)}, dzb_[3] = {T(0), T(0), T(0)}, d2zb_[3] = {T(0), T(0), T(0)};
    __device__ __forceinline__ T bfull(int k) const { return b0_[k] + u0[(3 + k) * SHP]; }
    __device__ __forceinline__ T dzb(int k) const { return dzb_[k]; }
    __device__ __forceinline__ T d2zb(int k) const { return d2zb_[k]; }
    __device__ __forceinline__ T hb() const { return hb_; }
    __device__ __forceinline__ T dzt0() const { return dzt0_; }
    __device__ __forceinline__ T d2zt0() const { return d2zt0_; }
    __device__ __forceinline__ T rodev() const { return rdev_; }
    __device__ __forceinline__ T ttfull() const { return t0_ + ws[3][1]; }
};

// T = double: the plain FP64 form, reading U (FP64) and bit-identical to the
//   reference's arithmetic.
// T = float: the mixed form, evaluating the fluxes in FP32 while the RK
//   registers stay FP64.  Stage 1 takes its stencils (rounded to FP32) and
//   Z = U from the FP64 state; stages 1-2 write the FP32 copy of their result
//   (the next stage's stencil input) and the FP64 Z; stage 3 reads the FP32
//   copy and Z and writes the FP64 state (and the FP64 CFL of the next step).
// WAVES: minimum waves per SIMD the register allocation must allow (2 keeps
// 256 VGPRs per wave; 1 lets the FP64 MHD variant use AGPRs instead of
// spilling).
// LDS of one work item: NRING planes of u, v, w [BX, BY, BZ] and NSTAR
// centre planes of RU, RV, RW, TT, P [RO].
template <typename T, bool MAG, bool ONE_BARRIER, bool RO_N = false>
struct StreamLds {
    static constexpr int NR = MAG ? 6 : 3;
    static constexpr int NS = RO_N ? 6 : 5;
    static constexpr int NRING = ONE_BARRIER ? 4 : 3;
    static constexpr int NSTAR = ONE_BARRIER ? 3 : 2;
    static constexpr int bytes = (int)sizeof(T) * (NRING * NR + NSTAR * NS) * SHP;
    // the mixed form's stage-1 ring of update operands (stage_item, ZLDS)
    // follows the planes, 16-byte aligned
    static constexpr int ring_at = (bytes + 15) & ~15;
};

// LDS of a kernel whose work items run stage_item<T, MAG', ..., STAGE,
// ONE_BARRIER, DEFER> for MAG' up to MAG: the planes, and in the mixed form's
// stage 1 with the deferred update and two barriers, two planes of FP64
// update operands per thread after them.
template <typename T, bool MAG, int STAGE, bool ONE_BARRIER, bool DEFER, bool RO_N = false>
struct StageLds {
    static constexpr bool ring = !std::is_same<T, double>::value && STAGE == 1 && DEFER && !ONE_BARRIER;
    static constexpr int bytes = ring ? StreamLds<T, MAG, ONE_BARRIER, RO_N>::ring_at +
                                            2 * (MAG ? NF_MAG : NF_HYDRO) * STX * STY * (int)sizeof(double)
                                      : StreamLds<T, MAG, ONE_BARRIER, RO_N>::bytes;
};

// One work item of a stage: the 32 x 8 tile (tile_x, tile_y) over the
// interior planes zb .. ze-1 (the wall planes are a separate launch: their
// one-sided stencils from global memory would raise this kernel's register
// budget).  lds: StageLds::bytes of shared memory.  Stage 3 accumulates the
// CFL terms of every point it updates into cfl.
// S32 (with T = double): the mixed form with the off-default branches --
// fluxes evaluated in FP64 (with the branches) from the FP32 copies of the
// stage values, which stages 1-2 write as the mixed form does.
template <typename T, bool MAG, bool ROT, int STAGE, bool ONE_BARRIER, bool DEFER, bool EXT = false,
          int EXTM = EXT_INTERIOR, bool S32 = false>
__device__ __forceinline__ void stage_item(const StageArgs& a, int tile_x, int tile_y, int zb,
                                           int ze, char* lds, CflAcc& cfl)
{
    constexpr bool F64 = std::is_same<T, double>::value;
    static_assert(F64 || !EXT, "the off-default branches are FP64 only");
    // RO of the centre plane in LDS: only for density diffusion (iD)
    constexpr bool RO_N = EXT && (EXTM & (EXT_DH | EXT_DV)) != 0;
    constexpr int NR = MAG ? 6 : 3;   // ring fields: u v w [BX BY BZ]
    constexpr int NS = RO_N ? 6 : 5;  // centre-plane fields: RU RV RW TT P [RO]
    constexpr int NF = MAG ? NF_MAG : NF_HYDRO;
    // ring depths: NRING planes of u, v, w (BX, BY, BZ) and NSTAR centre
    // planes, deep enough that one barrier per plane suffices (a thread can be
    // at most one plane ahead of another, and then it writes slots nobody
    // still reads)
    constexpr int NRING = ONE_BARRIER ? 4 : 3;
    constexpr int NSTAR = ONE_BARRIER ? 3 : 2;
    // Z-type operand of the update read from HBM: Z for stages 2-3, and in
    // the mixed form also the FP64 state for stage 1
    constexpr bool ZLOAD = STAGE > 1 || !F64;
    // storage type of the stencil input: the FP64 state, or the FP32 copy
    using TI = typename std::conditional<(F64 && !S32) || STAGE == 1, double, float>::type;
    T (*s_uvw)[NR][SHP] = reinterpret_cast<T (*)[NR][SHP]>(lds);
    T (*s_st)[NS][SHP] = reinterpret_cast<T (*)[NS][SHP]>(lds + sizeof(T) * NRING * NR * SHP);

    const Geom& g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int t = ty * STX + tx;

    const int x0 = tile_x * STX, y0 = tile_y * STY;

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

    const TI* __restrict__ U;
    if constexpr (std::is_same<TI, double>::value)
        U = a.Uin;
    else
        U = a.Uin32;
    const double* __restrict__ zsrc = (STAGE == 1) ? a.Uin : a.Z;
    // the Z operand of stages 2-3 is read once per point and by no other
    // tile: a nontemporal load keeps its lines from displacing the state's
    // halo lines in L2 (stage 1 reads it from the state itself)
    auto zld = [&](int64_t idx) -> double {
        if constexpr (STAGE > 1)
            return __builtin_nontemporal_load(zsrc + idx);
        else
            return zsrc[idx];
    };
    const int64_t fs = g.fs;
    // Mixed form, stage 1: the update's Z-type operand is the FP64 state at
    // the point, the very values the thread loaded for the stencil one
    // iteration before.  They wait in a thread-private LDS ring of two planes
    // instead of being read from HBM again (most of them have left L2 by
    // then: that read doubled the stage's fetched bytes).  Only with the
    // deferred update and two barriers, whose LDS leaves room for it.
    constexpr bool ZLDS = StageLds<T, MAG, STAGE, ONE_BARRIER, DEFER, RO_N>::ring;
    double* zring = reinterpret_cast<double*>(lds + StreamLds<T, MAG, ONE_BARRIER, RO_N>::ring_at);
    SAcc<T, MAG, RO_N> acc;
    T row[3] = {T(0), T(0), T(0)};
    T rinv[3] = {T(0), T(0), T(0)};
    T rdv[3] = {T(0), T(0), T(0)};
    // raw values of two planes, used in turn: the loop body for plane L
    // consumes one pair and loads plane L+1 into the other, so no value a
    // load is still writing is ever copied (a copy would wait for every
    // memory operation issued before it)
    RawPt<TI, MAG> rawA, haloA, rawB, haloB;
    load_raw<TI, MAG>(rawA, U, fs, (int64_t)(zb - 1) * g.plane + go);
    if (has_halo) load_raw<TI, MAG>(haloA, U, fs, (int64_t)(zb - 1) * g.plane + gh);

    const double DT = *a.dt;
    const double cg = a.c.GAM[STAGE - 1] * DT;
    const double cz = STAGE < 3 ? a.c.ZETA[STAGE - 1] * DT : 0.0;
    // Z-type operand of the next plane to update, loaded one iteration ahead
    double zpre[NF], zcur[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) zpre[f] = zcur[f] = 0.0;
    if (ZLOAD && !DEFER) {
#pragma unroll
        for (int f = 0; f < NF; ++f) zpre[f] = zld(f * fs + (int64_t)zb * g.plane + go);
    }

    // The column's x and y metric terms (the same on every plane) are loaded
    // once; the z terms of each plane, the same for the whole workgroup, by
    // scalar loads at the top of its iteration.
    const double mdxx = a.m.dxxdx[gx + 1];
    const double md2xx = a.m.d2xxdx2[gx + 1];
    const double mdyy = a.m.dyydy[gy + 1];
    const double md2yy = a.m.d2yydy2[gy + 1];
    // stage 3: the column's part of the TSTEP grid spacing (DDZ comes with
    // each plane's record)
    double ddxy = 0.0;
    if (STAGE == 3) {
        ddxy = a.m.DDY[gy + 1];
        if (a.c.xdim) ddxy = ref_min(ddxy, a.m.DDX[gx + 1]);
    }
    // RK update of plane zq from its tendencies o and its Z-type operand zin
    // (Uout = zin + GAM*DT*F, Z = Uout + ZETA*DT*F; stage 3 also the CFL terms)
    auto update = [&](int zq, const FluxOutT<T>& o, const double* zin) {
        if (!valid) return;
        const int64_t i = (int64_t)zq * g.plane + go;
        double un[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const double F = fsel(o, f);
            un[f] = zin[f] + cg * F;
            if ((F64 && !S32) || STAGE == 3)
                stream_store(&a.Uout[f * fs + i], un[f]);
            else
                stream_store(&a.Uout32[f * fs + i], float(un[f] - bg_sub(a, f, zq)));
            if (STAGE < 3) stream_store(&a.Z[f * fs + i], un[f] + cz * F);
        }
        if (STAGE == 3 && a.cfl_parts) {
            double wp[5], up;
            CflRec r;
            {
                // the plane's record, field by field through scalar loads
                double* dst = reinterpret_cast<double*>(&r);
                const double* src = reinterpret_cast<const double*>(a.m.cflz + zq);
#pragma unroll
                for (int k = 0; k < (int)(offsetof(CflRec, pad) / sizeof(double)); ++k)
                    dst[k] = uniform_load(src + k);
            }
            if constexpr (MAG)
                cfl_values_rec(r, true, un[0], un[1], un[2], un[3], un[4], un[5], un[6], un[7], ddxy,
                               wp, up);
            else
                cfl_values_rec(r, false, un[0], un[1], un[2], un[3], un[4], 0.0, 0.0, 0.0, ddxy, wp,
                               up);
            cfl.add(wp, up);
        }
    };
    // DEFER: the update of plane z runs one iteration later, after the next
    // plane is published, with its Z operand loaded right after the flux
    // evaluation (FP64: right before it) -- the operand's load latency then
    // overlaps the next plane's LDS work, and nothing but the 8 tendencies
    // and the operand is carried through it.  Otherwise the operand of plane z+1 is prefetched
    // one iteration ahead and plane z is updated at once.
    FluxOutT<T> fpend;
    double zpend[NF];
    int zp = -1;
#pragma unroll
    for (int f = 0; f < NF; ++f) zpend[f] = 0.0;
    // one plane L: derive and publish it from (own, halo), fetch plane L+1
    // into (own_n, halo_n), evaluate and update plane L-1
    auto plane = [&](int L, const RawPt<TI, MAG>& own, const RawPt<TI, MAG>& halo, RawPt<TI, MAG>& own_n,
                     RawPt<TI, MAG>& halo_n) __attribute__((always_inline)) {
        // z metric terms of plane L-1, whose fluxes this iteration evaluates
        // (when L-1 >= zb); L-1 >= zb-2 >= 0 is in range
        const double mdzz = uniform_load(a.m.dzzdz + (L - 1));
        const double md2zz = uniform_load(a.m.d2zzdz2 + (L - 1));
        const double mkap = uniform_load(a.m.RKAPA + (L - 1));
        const double mdkap = uniform_load(a.m.DKAPA + (L - 1));
        // -- derive plane L and publish it
        {
            const PrepPt<T> q = prep_point<T, MAG, TI>(own, a, L);
            if constexpr (ZLDS) {
                double* zr = &zring[(L & 1) * NF * STX * STY + t];
                zr[0 * STX * STY] = own.ru;
                zr[1 * STX * STY] = own.rv;
                zr[2 * STX * STY] = own.rw;
                zr[3 * STX * STY] = own.ro;
                zr[4 * STX * STY] = own.tt;
                if constexpr (MAG) {
                    zr[5 * STX * STY] = own.bx;
                    zr[6 * STX * STY] = own.by;
                    zr[7 * STX * STY] = own.bz;
                }
            }
            const T ri = T(1) / q.ro;
            const T u = q.ru * ri, v = q.rv * ri, w = q.rw * ri;
            T* su = &s_uvw[L % NRING][0][0];
            T* ss = &s_st[L % NSTAR][0][0];
            su[0 * SHP + pc] = u;
            su[1 * SHP + pc] = v;
            su[2 * SHP + pc] = w;
            ss[0 * SHP + pc] = q.ru;
            ss[1 * SHP + pc] = q.rv;
            ss[2 * SHP + pc] = q.rw;
            ss[3 * SHP + pc] = q.tt;
            ss[4 * SHP + pc] = q.p;
            if (RO_N) ss[5 * SHP + pc] = q.ro;
            if (MAG) {
                su[3 * SHP + pc] = q.bx;
                su[4 * SHP + pc] = q.by;
                su[5 * SHP + pc] = q.bz;
            }
            const T nw[5] = {q.ru, q.rv, q.rw, q.tt, q.p};
#pragma unroll
            for (int k = 0; k < SAcc<T, MAG, RO_N>::NW; ++k) {
                acc.ws[k][0] = acc.ws[k][1];
                acc.ws[k][1] = acc.ws[k][2];
                acc.ws[k][2] = nw[k];
            }
            row[0] = row[1]; row[1] = row[2]; row[2] = q.ro;
            rinv[0] = rinv[1]; rinv[1] = rinv[2]; rinv[2] = ri;
            if (!F64) { rdv[0] = rdv[1]; rdv[1] = rdv[2]; rdv[2] = q.rdev; }
            if (has_halo) {
                const PrepPt<T> h = prep_point<T, MAG, TI>(halo, a, L);
                const T hri = T(1) / h.ro;
                su[0 * SHP + ph] = h.ru * hri;
                su[1 * SHP + ph] = h.rv * hri;
                su[2 * SHP + ph] = h.rw * hri;
                ss[0 * SHP + ph] = h.ru;
                ss[1 * SHP + ph] = h.rv;
                ss[2 * SHP + ph] = h.rw;
                ss[3 * SHP + ph] = h.tt;
                ss[4 * SHP + ph] = h.p;
                if (RO_N) ss[5 * SHP + ph] = h.ro;
                if (MAG) {
                    su[3 * SHP + ph] = h.bx;
                    su[4 * SHP + ph] = h.by;
                    su[5 * SHP + ph] = h.bz;
                }
            }
        }
        // -- fetch plane L+1: EARLY_LOAD issues the loads as soon as plane L's
        // raw values are consumed, so they are in flight through the barrier
        // as well (measured faster in FP64 stages 2-3 and FP64 hydro; FP64
        // MHD stage 1 and the mixed form prefer them after the barrier)
        constexpr bool EARLY_LOAD = F64 && (STAGE > 1 || !MAG);
        if (EARLY_LOAD && L < ze) {
            load_raw<TI, MAG>(own_n, U, fs, (int64_t)(L + 1) * g.plane + go);
            if (has_halo) load_raw<TI, MAG>(halo_n, U, fs, (int64_t)(L + 1) * g.plane + gh);
        }
        lds_barrier();
        if (!EARLY_LOAD && L < ze) {
            load_raw<TI, MAG>(own_n, U, fs, (int64_t)(L + 1) * g.plane + go);
            if (has_halo) load_raw<TI, MAG>(halo_n, U, fs, (int64_t)(L + 1) * g.plane + gh);
        }
        if (DEFER && zp >= 0) {
            update(zp, fpend, zpend);
            zp = -1;
        }
        const int z = L - 1;
        // FP64 with DEFER: the Z operand of plane z is loaded before its flux
        // evaluation, whose duration then covers the load latency (measured
        // faster in FP64; in FP32 the 2-wave register budget prefers the load
        // after the flux)
        constexpr bool ZEARLY = DEFER && ZLOAD && F64;
        if (ZEARLY && z >= zb) {
            // straight into the pending operand (the update above consumed the
            // previous one): a copy of a value still in flight would wait for
            // every memory operation issued before its load
#pragma unroll
            for (int f = 0; f < NF; ++f) zpend[f] = zld(f * fs + (int64_t)z * g.plane + go);
        }
        if (!DEFER && ZLOAD && z >= zb) {
            // this iteration updates plane z; fetch the operand of plane z+1
#pragma unroll
            for (int f = 0; f < NF; ++f) zcur[f] = zpre[f];
            if (z + 1 < ze) {
#pragma unroll
                for (int f = 0; f < NF; ++f) zpre[f] = zld(f * fs + (int64_t)(z + 1) * g.plane + go);
            }
        }
        if (z >= zb) {
            acc.st = &s_st[z % NSTAR][0][pc];
            acc.um = &s_uvw[(z - 1) % NRING][0][pc];
            acc.u0 = &s_uvw[z % NRING][0][pc];
            acc.up = &s_uvw[(z + 1) % NRING][0][pc];
            acc.ro0 = row[1];
            acc.ri0 = rinv[1];
            if constexpr (RO_N) {
                acc.rz[0] = row[0];
                acc.rz[1] = row[1];
                acc.rz[2] = row[2];
            }
            if constexpr (!F64) {
                const int nz = g.nz;
                acc.rdev_ = rdv[1];
                acc.hb_ = uniform_load(a.bgf + BG_HB * nz + z);
                acc.dzt0_ = uniform_load(a.bgf + BG_DZT0 * nz + z);
                acc.d2zt0_ = uniform_load(a.bgf + BG_D2ZT0 * nz + z);
                acc.t0_ = uniform_load(a.bgf + BG_T0 * nz + z);
                if (MAG) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        acc.b0_[k] = uniform_load(a.bgf + (BG_B0 + k) * nz + z);
                        acc.dzb_[k] = uniform_load(a.bgf + (BG_DZB + k) * nz + z);
                        acc.d2zb_[k] = uniform_load(a.bgf + (BG_D2ZB + k) * nz + z);
                    }
                }
            }
            PointMetric pm;
            pm.dxx = mdxx;
            pm.d2xx = md2xx;
            pm.dyy = mdyy;
            pm.d2yy = md2yy;
            pm.dzz = mdzz;
            pm.d2zz = md2zz;
            pm.kap = mkap;
            pm.dkap = mdkap;
            FluxOutT<T> o;
            ExtPt e;
            if constexpr (EXT) ext_point<EXTM>(a, gx, gy, z, (int64_t)z * g.plane + go, e);
            flux_point<MAG, ROT, 0, T, !F64, EXT, EXTM>(acc, a.c, pm, o, &e);
            if constexpr (EXT && (EXTM & EXT_SHR)) {
                if (valid && ext_on<EXTM, EXT_SHR>(a.x.flags)) a.x.ft[(int64_t)z * g.plane + go] = o.FT;
            }
            // FP64 stage 1: Z = U, the centre values of plane z
            double uc[NF];
            if constexpr (!ZLOAD) {
                uc[0] = acc.ws[0][1];
                uc[1] = acc.ws[1][1];
                uc[2] = acc.ws[2][1];
                uc[3] = row[1];
                uc[4] = acc.ws[3][1];
                if constexpr (MAG) {
                    uc[5] = acc.u0[3 * SHP];
                    uc[6] = acc.u0[4 * SHP];
                    uc[7] = acc.u0[5 * SHP];
                }
            }
            if constexpr (DEFER) {
                fpend = o;
                zp = z;
                if constexpr (ZLDS) {
                    // plane z's values, stored by this thread when it derived plane z
                    const double* zr = &zring[(z & 1) * NF * STX * STY + t];
#pragma unroll
                    for (int f = 0; f < NF; ++f) zpend[f] = zr[f * STX * STY];
                } else if constexpr (!ZEARLY) {
#pragma unroll
                    for (int f = 0; f < NF; ++f)
                        zpend[f] = ZLOAD ? zld(f * fs + (int64_t)z * g.plane + go) : uc[f];
                }
            } else {
                update(z, o, ZLOAD ? zcur : uc);
            }
        }
        if (!ONE_BARRIER) lds_barrier();
    };
    // Everything loaded before the loop (the first plane, the column's metric
    // terms, DT) is waited for here, once.  Otherwise the wait-count pass
    // merges those loads into the loop header's state and makes iterations
    // wait for them with counts that also drain the next plane's loads.
    __builtin_amdgcn_s_waitcnt(0);
    for (int L = zb - 1; L <= ze; L += 2) {
        plane(L, rawA, haloA, rawB, haloB);
        if (L + 1 > ze) break;
        plane(L + 1, rawB, haloB, rawA, haloA);
    }
    if (DEFER && zp >= 0) update(zp, fpend, zpend);
}

// One stage of one domain: block -> (tile, z chunk) work item.  With
// xcd_remap, consecutive tiles go to the same XCD (blocks are dealt to the 8
// XCDs round-robin) so halo re-reads hit its L2.  Chunk k covers the planes
// 2 + k*zchunk .. (the last chunk runs to nPz-1).
template <typename T, bool MAG, bool ROT, int STAGE, int WAVES, bool ONE_BARRIER, bool DEFER,
          bool EXT = false, int EXTM = EXT_INTERIOR, bool S32 = false>
__global__ void __launch_bounds__(256, WAVES) stage_stream_kernel(StageArgs a, int zchunk, int nch,
                                                                  int ntx, int nty, int xcd_remap)
{
    __shared__ __align__(16)
        char lds[StageLds<T, MAG, STAGE, ONE_BARRIER, DEFER, EXT && (EXTM & (EXT_DH | EXT_DV)) != 0>::bytes];
    if (domain_halted(a)) return;
    const int nb = gridDim.x;
    const int bid = blockIdx.x;
    const int logical = xcd_remap ? (bid % 8) * (nb / 8) + bid / 8 : bid;
    const int tile_x = logical % ntx;
    const int rest = logical / ntx;
    const int tile_y = rest % nty;
    const int chunk = rest / nty;
    const int zb = 2 + chunk * zchunk;
    const int ze = chunk == nch - 1 ? a.g.nPz : zb + zchunk;
    CflAcc cfl;   // stage 3: TSTEP of the next step on the new state
    stage_item<T, MAG, ROT, STAGE, ONE_BARRIER, DEFER, EXT, EXTM, S32>(a, tile_x, tile_y, zb, ze, lds, cfl);
    if (STAGE == 3 && a.cfl_parts) {
        // the reduction borrows the plane buffers (a separate 768 B would
        // take an iD variant's two blocks per CU over the CU's LDS)
        __syncthreads();
        cfl_block_reduce(cfl, a.cfl_parts, a.cfl_base + blockIdx.x, reinterpret_cast<double (*)[16]>(lds));
    }
}

struct StreamPlan {
    int zchunk, nchunks, ntx, nty, nblocks, xcd_remap;
};

// Work decomposition of one domain: 32 x 8 tiles times z chunks of at least
// 8 planes (the last chunk takes the remainder), about target_blocks items.
// The launch is fastest when its blocks fill the device's resident slots
// once (no second round, no chunk-overlap planes beyond what that needs), so
// the caller passes the resident block count of the kernel variant.
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
    p.zchunk = nint / nch;
    p.nchunks = nch;
    p.nblocks = tiles * p.nchunks;
    p.xcd_remap = (p.nblocks % 8 == 0) ? 1 : 0;
    return p;
}

// Launch of one interior stage.  sched bit 0: one barrier per plane
// (ONE_BARRIER), bit 1: deferred update (DEFER); waves: register budget.
template <typename T, bool MAG, bool ROT, int STAGE, int WAVES>
static void launch_stream_w(const StageArgs& a, hipStream_t s, const StreamPlan& p, int sched)
{
    dim3 blk(STX, STY, 1);
    switch (sched & 3) {
    case 0:
        stage_stream_kernel<T, MAG, ROT, STAGE, WAVES, false, false><<<p.nblocks, blk, 0, s>>>(
            a, p.zchunk, p.nchunks, p.ntx, p.nty, p.xcd_remap);
        break;
    case 1:
        stage_stream_kernel<T, MAG, ROT, STAGE, WAVES, true, false><<<p.nblocks, blk, 0, s>>>(
            a, p.zchunk, p.nchunks, p.ntx, p.nty, p.xcd_remap);
        break;
    case 2:
        stage_stream_kernel<T, MAG, ROT, STAGE, WAVES, false, true><<<p.nblocks, blk, 0, s>>>(
            a, p.zchunk, p.nchunks, p.ntx, p.nty, p.xcd_remap);
        break;
    default:
        stage_stream_kernel<T, MAG, ROT, STAGE, WAVES, true, true><<<p.nblocks, blk, 0, s>>>(
            a, p.zchunk, p.nchunks, p.ntx, p.nty, p.xcd_remap);
        break;
    }
}

// Kernel function of a variant (stage 2 stands for all stages: they share
// the resource footprint up to the stage-3 CFL partials).
template <typename T, bool MAG, bool ROT>
static const void* stream_fn(int waves, int sched)
{
    if (waves == 1) {
        switch (sched & 3) {
        case 0: return (const void*)stage_stream_kernel<T, MAG, ROT, 2, 1, false, false>;
        case 1: return (const void*)stage_stream_kernel<T, MAG, ROT, 2, 1, true, false>;
        case 2: return (const void*)stage_stream_kernel<T, MAG, ROT, 2, 1, false, true>;
        default: return (const void*)stage_stream_kernel<T, MAG, ROT, 2, 1, true, true>;
        }
    }
    switch (sched & 3) {
    case 0: return (const void*)stage_stream_kernel<T, MAG, ROT, 2, 2, false, false>;
    case 1: return (const void*)stage_stream_kernel<T, MAG, ROT, 2, 2, true, false>;
    case 2: return (const void*)stage_stream_kernel<T, MAG, ROT, 2, 2, false, true>;
    default: return (const void*)stage_stream_kernel<T, MAG, ROT, 2, 2, true, true>;
    }
}

// Blocks of a variant resident on the whole device at once.
inline int resident_blocks(const void* fn, int threads)
{
    int per_cu = 0, dev = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 256;
    return per_cu * ncu;
}

template <typename T, bool MAG, bool ROT, int STAGE>
static void launch_stream_stage(const StageArgs& a, hipStream_t s, int waves, int sched,
                                int target_blocks)
{
    const StreamPlan p = stream_plan(a.g, target_blocks);
    if (waves == 1)
        launch_stream_w<T, MAG, ROT, STAGE, 1>(a, s, p, sched);
    else
        launch_stream_w<T, MAG, ROT, STAGE, 2>(a, s, p, sched);
}

}  // namespace mhd
