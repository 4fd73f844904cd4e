// fluxes() of the reference (3dmhdsub.h:194-696) evaluated at one point.
//
// The reference builds each flux with whole-array sweeps (about 73 for hydro,
// 155 for MHD with rotation); here every flux of one point is accumulated in
// registers, in the same statement order and with the same left-to-right
// products as the operator templates of 3dmhdutils.h:195-239, so each value is
// rounded exactly as in the reference (the file is compiled without
// floating-point contraction).  Intermediate arrays of the reference (WW1..3
// as D1x(u), D1z(w), ...; RU..TT as lMag scratch) become locals; the mixed
// derivatives D1y(D1x u) etc. evaluate their inner derivative at the
// neighbours the outer one needs, exactly as the reference's extended-range
// sweeps do (3dmhdsub.h:475-476,505,520-521).
//
// Accessor A provides, for offsets (dx,dy,dz) from the point:
//   ru rv rw ro tt bx by bz  state values
//   ri = 1/RO, u = RU*ri, v = RV*ri, w = RW*ri, p = RO*TT   (3dmhdsub.h:267-276)
// WALL = 0 (interior plane), 1 (top wall, z = zl), 2 (bottom wall, z = zr-1):
// the one-sided stencils of 3dmhdsub.h:211-226, 572-584, 640-654, 673-687.
//
// DEV = true (the mixed form only): the accessor's ro/tt/p stencil values are
// deviations from a fixed z-only background RO0(z), TT0(z), P0 = RO0*TT0, so
// FP32 differences of a nearly hydrostatic state do not cancel.  x and y
// differences of the background vanish; its z differences come from the
// accessor, precomputed in FP64: dzt0() = D1z(TT0), d2zt0() = D2z(TT0) and
// hb() = RO0*GRAV - D1z(P0) (the hydrostatic residual).  rodev() is the
// centre density deviation and ttfull() the full centre temperature.  BX, BY,
// BZ stencil values are deviations from BX0(z), BY0(z), BZ0(z) likewise, with
// bfull(k) the full centre field and dzb(k), d2zb(k) the background's D1z and
// D2z (one-sided at the walls for BX and BY, as the stencils below).
#pragma once
// FP64 arithmetic follows the reference's operation order, so nothing is
// contracted into FMA; the translation unit of the FP32 (mixed-form) kernels
// alone allows contraction within an expression.
#ifdef MHD3D_FP32_FORM_TU
#pragma clang fp contract(on)
#else
#pragma clang fp contract(off)
#endif

#include <type_traits>

#include "mhd3d_types.hpp"

namespace mhd {

// The eight tendencies of one point, in the type the fluxes are evaluated in.
template <typename T>
struct FluxOutT {
    T FU, FV, FW, FR, FT, B1, B2, B3;
};
using FluxOut = FluxOutT<double>;

struct PointMetric {
    double dxx, d2xx, dyy, d2yy, dzz, d2zz, kap, dkap;
};

template <typename T>
struct Ops {
    const T dxx, d2xx, dyy, d2yy, dzz, d2zz, hx, h2x, hy, h2y, hz, h2z;
    // func_d1: ((f[+1]-f[-1]) * du) * h     (3dmhdutils.h:195-199)
    __device__ __forceinline__ T d1x(T fp, T fm) const { return ((fp - fm) * dxx) * hx; }
    __device__ __forceinline__ T d1y(T fp, T fm) const { return ((fp - fm) * dyy) * hy; }
    __device__ __forceinline__ T d1z(T fp, T fm) const { return ((fp - fm) * dzz) * hz; }
    // func_d2: (df*d2u*h + ((f[+1]-2f)+f[-1])*du^2*h2)   (3dmhdutils.h:215-220)
    __device__ __forceinline__ T d2x(T fp, T f0, T fm) const
    {
        return (fp - fm) * d2xx * hx + ((fp - T(2) * f0) + fm) * (dxx * dxx) * h2x;
    }
    __device__ __forceinline__ T d2y(T fp, T f0, T fm) const
    {
        return (fp - fm) * d2yy * hy + ((fp - T(2) * f0) + fm) * (dyy * dyy) * h2y;
    }
    __device__ __forceinline__ T d2z(T fp, T f0, T fm) const
    {
        return (fp - fm) * d2zz * hz + ((fp - T(2) * f0) + fm) * (dzz * dzz) * h2z;
    }
};

template <typename T>
__device__ __forceinline__ Ops<T> make_ops(const Phys& c, const PointMetric& m)
{
    return Ops<T>{T(m.dxx), T(m.d2xx), T(m.dyy), T(m.d2yy), T(m.dzz), T(m.d2zz),
                  T(c.hx),  T(c.h2x),  T(c.hy),  T(c.h2y),  T(c.hz),  T(c.h2z)};
}

// Contributions of the magnetic block to FU, FV, FW, FT, in the order the
// reference adds them (3dmhdsub.h:592-690).  The ops a sink receives, in
// sequence (U/V/W/T = FU/FV/FW/FT, + add, - subtract):
//   U+ U- U- V+ V- W+ T+ T+ T+ T+ | U+ W- T+ T+ | V+ W- T+ T+ T+
enum MagOp { MAG_UA, MAG_US, MAG_VA, MAG_VS, MAG_WA, MAG_WS, MAG_TA };
constexpr int NMAGTERM = 19;
constexpr MagOp kMagOps[NMAGTERM] = {MAG_UA, MAG_US, MAG_US, MAG_VA, MAG_VS, MAG_WA, MAG_TA,
                                     MAG_TA, MAG_TA, MAG_TA, MAG_UA, MAG_WS, MAG_TA, MAG_TA,
                                     MAG_VA, MAG_WS, MAG_TA, MAG_TA, MAG_TA};

// Applies one contribution to the accumulators.
template <typename T>
__device__ __forceinline__ void mag_apply(MagOp op, T v, T& FU, T& FV, T& FW, T& FT)
{
    switch (op) {
    case MAG_UA: FU += v; break;
    case MAG_US: FU -= v; break;
    case MAG_VA: FV += v; break;
    case MAG_VS: FV -= v; break;
    case MAG_WA: FW += v; break;
    case MAG_WS: FW -= v; break;
    default: FT += v; break;
    }
}

// Sink that applies each contribution at once (the single-thread flux).
template <typename T>
struct InlineSink {
    T &FU, &FV, &FW, &FT;
    __device__ __forceinline__ void put(MagOp op, T v) { mag_apply(op, v, FU, FV, FW, FT); }
};

// Inputs of the off-default branches at one point (FP64 form only), formed
// by the kernel from ExtArgs: the horizontal means of the neighbouring
// planes, the plane's density-diffusion weight and mean-profile conduction,
// the carried-over FT (lShr) and the heat-loss factors.
struct ExtPt {
    int flags = 0;
    double rom[3] = {0, 0, 0};   // ROM at z-1, z, z+1
    double ttm[3] = {0, 0, 0};   // TTM at z-1, z, z+1
    double wwz = 0;              // density-diffusion weight of the plane (:244-252)
    double cdh = 0, cdv = 0;     // DH*ORE, DV*ORE
    double hrad = 0;             // HRAD of the plane (:322-325)
    double ft_old = 0;           // lShr: FT before this call
    double heat_amp = 0, hx = 0, hy = 0, hz = 0;   // iTCon 2 heat loss
    double corr = 0, rscale = 0;                   // RLAX: CORRECT(z), RLAX * TFAC
};

// Hydrodynamic part of fluxes(): continuity, pressure and gravity,
// advection, conduction, viscous terms, viscous heating, pressure work and
// Coriolis (3dmhdsub.h:204-545).
//
// EXT = true adds the off-default branches selected by e->flags (FP64):
// density diffusion of W = RO - ROM(z) into FR (iD; compiled only for an
// accessor that gives RO at the neighbours, A::kNeighbourRO), the lRem conduction of
// TT - TTM(z), the carried-over FT of lShr (its advection, viscous-heating
// and pressure-work terms are skipped, :293,477-538), the heat loss of
// iTCon 2 and the RLAX correction, each at its place in the reference's
// statement order.  EXTM (a mask of EXT_* bits) limits the branches compiled
// in; the flags of e must lie within it.
template <bool ROT, int WALL, typename T, bool DEV, typename A, bool EXT = false,
          int EXTM = EXT_INTERIOR>
__device__ __forceinline__ void flux_hydro(const A& a, const Phys& c, const PointMetric& m, T& FU,
                                           T& FV, T& FW, T& FR, T& FT, const ExtPt* e = nullptr)
{
    // lShr drops every FT term inside the reference's `#if !lShr` blocks
    const bool heat_terms = !EXT || !ext_on<EXTM, EXT_SHR>(e->flags);
    const Ops<T> op = make_ops<T>(c, m);
    const T ORE = T(c.ORE), OCV = T(c.OCV);
    const T ORE_C43 = T(c.ORE_C43);
    const T C13_ORE = T(c.C13_ORE);
    T tmp;

    // ---- continuity :205-227
    FR = op.d1x(a.ru(1, 0, 0), a.ru(-1, 0, 0)) * T(-1);
    FR -= op.d1y(a.rv(0, 1, 0), a.rv(0, -1, 0));
    T dzrw;
    if (WALL == 1)
        dzrw = (T(4) * a.rw(0, 0, 1) - a.rw(0, 0, 2)) * (T(c.hz) * op.dzz);
    else if (WALL == 2)
        dzrw = (a.rw(0, 0, -2) - T(4) * a.rw(0, 0, -1)) * (T(c.hz) * op.dzz);
    else
        dzrw = op.d1z(a.rw(0, 0, 1), a.rw(0, 0, -1));
    FR -= dzrw;

    // ---- density diffusion :229-264 (func_d2 of W = RO - ROM(z), then
    // func_mul_extend: (d2 * wwz[z]) * (DH*ORE)).  The reference first copies
    // the planes two in onto the z ghost planes of RO (:230-239), so on a
    // wall W beyond the wall equals W two planes in.
    if constexpr (EXT && (EXTM & (EXT_DH | EXT_DV)) && A::kNeighbourRO) {
        if (e->flags & (EXT_DH | EXT_DV)) {
            const double r0 = e->rom[1];
            const double w0 = a.ro(0, 0, 0) - r0;
            if (e->flags & EXT_DH) {
                FR += (op.d2x(a.ro(1, 0, 0) - r0, w0, a.ro(-1, 0, 0) - r0) * e->wwz) * e->cdh;
                FR += (op.d2y(a.ro(0, 1, 0) - r0, w0, a.ro(0, -1, 0) - r0) * e->wwz) * e->cdh;
            }
            if (e->flags & EXT_DV) {
                double wp, wm;
                if (WALL == 1) {
                    wp = a.ro(0, 0, 1) - e->rom[2];
                    wm = wp;
                } else if (WALL == 2) {
                    wm = a.ro(0, 0, -1) - e->rom[0];
                    wp = wm;
                } else {
                    wp = a.ro(0, 0, 1) - e->rom[2];
                    wm = a.ro(0, 0, -1) - e->rom[0];
                }
                FR += (op.d2z(wp, w0, wm) * e->wwz) * e->cdv;
            }
        }
    }

    // ---- gravity and pressure :266-272
    if constexpr (DEV)
        FW = a.hb() + a.rodev() * T(c.GRAV);
    else
        FW = a.ro(0, 0, 0) * T(c.GRAV);
    FW -= op.d1z(a.p(0, 0, 1), a.p(0, 0, -1));
    FV = op.d1y(a.p(0, 1, 0), a.p(0, -1, 0)) * T(-1);
    FU = op.d1x(a.p(1, 0, 0), a.p(-1, 0, 0)) * T(-1);

    // ---- advection of momentum :279-289 (func_mul_d1: ((F+G+ - F-G-)*du)*h)
    FU -= ((a.ru(1, 0, 0) * a.u(1, 0, 0) - a.ru(-1, 0, 0) * a.u(-1, 0, 0)) * op.dxx) * op.hx;
    FU -= ((a.ru(0, 1, 0) * a.v(0, 1, 0) - a.ru(0, -1, 0) * a.v(0, -1, 0)) * op.dyy) * op.hy;
    FU -= ((a.ru(0, 0, 1) * a.w(0, 0, 1) - a.ru(0, 0, -1) * a.w(0, 0, -1)) * op.dzz) * op.hz;
    FV -= ((a.rv(1, 0, 0) * a.u(1, 0, 0) - a.rv(-1, 0, 0) * a.u(-1, 0, 0)) * op.dxx) * op.hx;
    FV -= ((a.rv(0, 1, 0) * a.v(0, 1, 0) - a.rv(0, -1, 0) * a.v(0, -1, 0)) * op.dyy) * op.hy;
    FV -= ((a.rv(0, 0, 1) * a.w(0, 0, 1) - a.rv(0, 0, -1) * a.w(0, 0, -1)) * op.dzz) * op.hz;
    FW -= ((a.rw(1, 0, 0) * a.u(1, 0, 0) - a.rw(-1, 0, 0) * a.u(-1, 0, 0)) * op.dxx) * op.hx;
    FW -= ((a.rw(0, 1, 0) * a.v(0, 1, 0) - a.rw(0, -1, 0) * a.v(0, -1, 0)) * op.dyy) * op.hy;
    FW -= ((a.rw(0, 0, 1) * a.w(0, 0, 1) - a.rw(0, 0, -1) * a.w(0, 0, -1)) * op.dzz) * op.hz;

    const T RI = a.ri(0, 0, 0);
    const T U0 = a.u(0, 0, 0), V0 = a.v(0, 0, 0), W0 = a.w(0, 0, 0);
    const T T0 = a.tt(0, 0, 0);   // stencil value (a deviation when DEV)
    T T0f = T0;                    // full temperature, for products
    if constexpr (DEV) T0f = a.ttfull();

    // ---- temperature advection :294-296 (func_d1_mul: (((df*du)*h)*g)*s)
    const T dxT = op.d1x(a.tt(1, 0, 0), a.tt(-1, 0, 0));
    const T dyT = op.d1y(a.tt(0, 1, 0), a.tt(0, -1, 0));
    T dzT = op.d1z(a.tt(0, 0, 1), a.tt(0, 0, -1));
    if constexpr (DEV) dzT += a.dzt0();
    if (heat_terms) {
        FT = (dxT * U0) * T(-1);
        FT -= dyT * V0;
        FT -= dzT * W0;
    } else {
        FT = e->ft_old;   // lShr: FT is never reset (the SET above is inside `#if !lShr`)
    }

    // ---- conduction :315-319 (func_d2_mul_array_z: ((d2*g)*a[z])*s)
    tmp = T(c.OCV_REPR);
    if (EXT && ext_on<EXTM, EXT_REM>(e->flags)) {
        if (ext_on<EXTM, EXT_SHR>(e->flags)) {
            // lRem with lShr: the TTM update and WW1 = TT - TTM sit inside
            // `#if !lShr` (:297-302), so WW1 still holds P = RO*TT (:267)
            // and the static TTM stays zero, which makes HRAD zero
            const T p0 = a.p(0, 0, 0);
            FT += (op.d2x(a.p(1, 0, 0), p0, a.p(-1, 0, 0)) * RI) * tmp;
            FT += (op.d2y(a.p(0, 1, 0), p0, a.p(0, -1, 0)) * RI) * tmp;
            FT += (op.d2z(a.p(0, 0, 1), p0, a.p(0, 0, -1)) * RI) * tmp;
            FT += (RI * T(0)) * tmp;
        } else {
            // lRem :304-313: func_d2_mul of WW1 = TT - TTM(z) ((d2 * RO) * s, RO
            // holding 1/RO), then func_mul_extend of HRAD: (RO * HRAD[z]) * s
            const T m0 = e->ttm[1];
            const T t0 = T0 - m0;
            FT += (op.d2x(a.tt(1, 0, 0) - m0, t0, a.tt(-1, 0, 0) - m0) * RI) * tmp;
            FT += (op.d2y(a.tt(0, 1, 0) - m0, t0, a.tt(0, -1, 0) - m0) * RI) * tmp;
            FT += (op.d2z(a.tt(0, 0, 1) - e->ttm[2], t0, a.tt(0, 0, -1) - e->ttm[0]) * RI) * tmp;
            FT += (RI * e->hrad) * tmp;
        }
    } else {
        const T kap = T(m.kap);
        FT += ((op.d2x(a.tt(1, 0, 0), T0, a.tt(-1, 0, 0)) * RI) * kap) * tmp;
        FT += ((op.d2y(a.tt(0, 1, 0), T0, a.tt(0, -1, 0)) * RI) * kap) * tmp;
        T d2zT = op.d2z(a.tt(0, 0, 1), T0, a.tt(0, 0, -1));
        if constexpr (DEV) d2zT += a.d2zt0();
        FT += ((d2zT * RI) * kap) * tmp;
        FT += ((dzT * RI) * T(m.dkap)) * tmp;
    }
    // ---- RLAX relaxation :426 (func_base_extend: FT += CORRECT[z] * (RLAX*TFAC))
    if (EXT && ext_on<EXTM, EXT_RLAX>(e->flags)) FT += e->corr * e->rscale;
    // ---- embedded heat loss, iTCon 2 :443-461: FT -= (((amp*RO)*gx)*gy)*gz
    if (EXT && ext_on<EXTM, EXT_HEAT>(e->flags))
        FT -= (((e->heat_amp * RI) * e->hx) * e->hy) * e->hz;

    // ---- viscous :462-472
    FU += op.d2x(a.u(1, 0, 0), U0, a.u(-1, 0, 0)) * ORE_C43;
    FU += op.d2y(a.u(0, 1, 0), U0, a.u(0, -1, 0)) * ORE;
    FU += op.d2z(a.u(0, 0, 1), U0, a.u(0, 0, -1)) * ORE;
    FV += op.d2x(a.v(1, 0, 0), V0, a.v(-1, 0, 0)) * ORE;
    FV += op.d2y(a.v(0, 1, 0), V0, a.v(0, -1, 0)) * ORE_C43;
    FV += op.d2z(a.v(0, 0, 1), V0, a.v(0, 0, -1)) * ORE;
    FW += op.d2x(a.w(1, 0, 0), W0, a.w(-1, 0, 0)) * ORE;
    FW += op.d2y(a.w(0, 1, 0), W0, a.w(0, -1, 0)) * ORE;
    FW += op.d2z(a.w(0, 0, 1), W0, a.w(0, 0, -1)) * ORE_C43;

    // ---- cross-viscous terms, viscous heating, pressure work :475-538
    const T dxu = op.d1x(a.u(1, 0, 0), a.u(-1, 0, 0));
    const T dxv = op.d1x(a.v(1, 0, 0), a.v(-1, 0, 0));
    const T dxw = op.d1x(a.w(1, 0, 0), a.w(-1, 0, 0));
    if (heat_terms) {
        tmp = T(c.OCV_ORE);
        FT += ((RI * dxu) * dxu) * T(c.OCV_ORE_C43);
        FT += ((RI * dxv) * dxv) * tmp;
        FT += ((RI * dxw) * dxw) * tmp;
        FT -= (dxu * T0f) * OCV;
    }
    {
        // D1x(v), D1x(u) at y+-1 (WW2, WW1 over FORRANGE_EXT_Y, :475-476)
        const T dxv_p = op.d1x(a.v(1, 1, 0), a.v(-1, 1, 0));
        const T dxv_m = op.d1x(a.v(1, -1, 0), a.v(-1, -1, 0));
        const T dxu_p = op.d1x(a.u(1, 1, 0), a.u(-1, 1, 0));
        const T dxu_m = op.d1x(a.u(1, -1, 0), a.u(-1, -1, 0));
        FU += op.d1y(dxv_p, dxv_m) * C13_ORE;
        FV += op.d1y(dxu_p, dxu_m) * C13_ORE;
    }
    const T dyu = op.d1y(a.u(0, 1, 0), a.u(0, -1, 0));
    const T dyv = op.d1y(a.v(0, 1, 0), a.v(0, -1, 0));
    if (heat_terms) {
        tmp = T(c.OCV_ORE);
        FT += ((RI * dyu) * dyu) * tmp;
        FT += ((RI * dyu) * dxv) * T(c.OCV_ORE_2);
        tmp = T(c.C43_OCV_ORE);
        FT += ((RI * dyv) * dyv) * tmp;
        FT -= (dyv * T0f) * OCV;
    }
    // D1z(w) at the point and at x+-1, y+-1 (WW3 over FORRANGE_EXT_XY, :505)
    const T dzw = op.d1z(a.w(0, 0, 1), a.w(0, 0, -1));
    if (heat_terms) {
        tmp = T(c.C43_OCV_ORE);
        FT += ((RI * dzw) * dzw) * tmp;
        FT += ((RI * dxu) * dzw) * (-tmp);
        FT += ((RI * dxu) * dyv) * (-tmp);
        FT += ((RI * dyv) * dzw) * (-tmp);
        FT -= (dzw * T0f) * OCV;
    }
    {
        const T dzw_xp = op.d1z(a.w(1, 0, 1), a.w(1, 0, -1));
        const T dzw_xm = op.d1z(a.w(-1, 0, 1), a.w(-1, 0, -1));
        const T dzw_yp = op.d1z(a.w(0, 1, 1), a.w(0, 1, -1));
        const T dzw_ym = op.d1z(a.w(0, -1, 1), a.w(0, -1, -1));
        FU += op.d1x(dzw_xp, dzw_xm) * C13_ORE;
        FV += op.d1y(dzw_yp, dzw_ym) * C13_ORE;
    }
    // D1z(u) at x+-1 (WW1 over FORRANGE_EXT_X) and D1z(v) at y+-1 (WW2, :520-521)
    const T dzu = op.d1z(a.u(0, 0, 1), a.u(0, 0, -1));
    const T dzv = op.d1z(a.v(0, 0, 1), a.v(0, 0, -1));
    const T dyw = op.d1y(a.w(0, 1, 0), a.w(0, -1, 0));
    if (heat_terms) {
        tmp = T(c.OCV_ORE);
        FT += ((RI * dzu) * dzu) * tmp;
        FT += ((RI * dzv) * dzv) * tmp;
        FT += ((RI * dyw) * dyw) * tmp;
        FT += ((RI * dzv) * dyw) * T(c.OCV_ORE_2);
    }
    {
        const T dzu_xp = op.d1z(a.u(1, 0, 1), a.u(1, 0, -1));
        const T dzu_xm = op.d1z(a.u(-1, 0, 1), a.u(-1, 0, -1));
        const T dzv_yp = op.d1z(a.v(0, 1, 1), a.v(0, 1, -1));
        const T dzv_ym = op.d1z(a.v(0, -1, 1), a.v(0, -1, -1));
        FW += op.d1x(dzu_xp, dzu_xm) * C13_ORE;
        FW += op.d1y(dzv_yp, dzv_ym) * C13_ORE;
    }
    if (heat_terms) {
        FT += ((RI * dzu) * dxw) * T(c.OCV_ORE_2);
    }

    // ---- Coriolis :540-545
    if (ROT) {
        FU += a.rv(0, 0, 0) * T(c.OMZ);
        FV -= a.ru(0, 0, 0) * T(c.OMZ);
        FV += a.rw(0, 0, 0) * T(c.OMX);
        FW -= a.rv(0, 0, 0) * T(c.OMX);
    }

}

// Magnetic part of fluxes(): induction (B1, B2, B3 = WW1..WW3), Lorentz
// force and Joule heating (3dmhdsub.h:547-695); the contributions to FU, FV,
// FW, FT go to the sink in the reference's order (kMagOps).
template <int WALL, typename T, bool DEV, typename A, typename S>
__device__ __forceinline__ void flux_mag(const A& a, const Phys& c, const PointMetric& m, S& sink,
                                         T& B1, T& B2, T& B3)
{
    const Ops<T> op = make_ops<T>(c, m);
    const T RI = a.ri(0, 0, 0);
    {
        // The reference recomputes D1x/D1y/D1z of u, v, w here (:548-570).
        // In FP32 (two waves per SIMD) a compiler memory fence forces the
        // recomputation instead of keeping the hydro part's values live;
        // the FP64 kernel runs one wave per SIMD and has the accumulation
        // registers to keep them (measured faster without the fence).  Same
        // operands, same results either way.
        if constexpr (!std::is_same<T, double>::value) asm volatile("" ::: "memory");
        const T dxu = op.d1x(a.u(1, 0, 0), a.u(-1, 0, 0));
        const T dxv = op.d1x(a.v(1, 0, 0), a.v(-1, 0, 0));
        const T dxw = op.d1x(a.w(1, 0, 0), a.w(-1, 0, 0));
        const T dyu = op.d1y(a.u(0, 1, 0), a.u(0, -1, 0));
        const T dyv = op.d1y(a.v(0, 1, 0), a.v(0, -1, 0));
        const T dyw = op.d1y(a.w(0, 1, 0), a.w(0, -1, 0));
        const T dzu = op.d1z(a.u(0, 0, 1), a.u(0, 0, -1));
        const T dzv = op.d1z(a.v(0, 0, 1), a.v(0, 0, -1));
        const T dzw = op.d1z(a.w(0, 0, 1), a.w(0, 0, -1));
        const T U0 = a.u(0, 0, 0), V0 = a.v(0, 0, 0), W0 = a.w(0, 0, 0);
        const T OBETA = T(c.OBETA), ORM = T(c.ORM);
        const T KJ = T(c.KJ);
        const T KJ2 = T(c.KJ2);
        // centre stencil values (deviations when DEV) and full fields for products
        const T BXs = a.bx(0, 0, 0), BYs = a.by(0, 0, 0), BZs = a.bz(0, 0, 0);
        T BX = BXs, BY = BYs, BZ = BZs;
        if constexpr (DEV) {
            BX = a.bfull(0);
            BY = a.bfull(1);
            BZ = a.bfull(2);
        }
        T s1, s2, s3, s4;
        // :548-556 (RU,RV,RW = D1x of u,v,w)
        B2 = BX * dxv;
        B2 -= BY * dxu;
        B3 = BX * dxw;
        B3 -= BZ * dxu;
        // :558-566
        const T dyw_ = dyw;
        B1 = BY * dyu;
        B1 -= BX * dyv;
        B3 += BY * dyw_;
        B3 -= BZ * dyv;
        // :568-590, one-sided D1z(w) at the walls
        T dzw_m;
        if (WALL == 1)
            dzw_m = (T(4) * a.w(0, 0, 1) - a.w(0, 0, 2)) * T(c.hz) * op.dzz;
        else if (WALL == 2)
            dzw_m = (T(-4) * a.w(0, 0, -1) + a.w(0, 0, -2)) * T(c.hz) * op.dzz;
        else
            dzw_m = dzw;
        B1 += BZ * dzu;
        B1 -= BX * dzw_m;
        B2 += BZ * dzv;
        B2 -= BY * dzw_m;
        // :592-607
        s1 = op.d1x(a.bx(1, 0, 0), a.bx(-1, 0, 0));
        s2 = op.d1x(a.by(1, 0, 0), a.by(-1, 0, 0));
        s3 = op.d1x(a.bz(1, 0, 0), a.bz(-1, 0, 0));
        s4 = op.d1y(a.bx(0, 1, 0), a.bx(0, -1, 0));
        sink.put(MAG_UA, (BY * s4) * OBETA);
        sink.put(MAG_US, (BY * s2) * OBETA);
        sink.put(MAG_US, (BZ * s3) * OBETA);
        sink.put(MAG_VA, (BX * s2) * OBETA);
        sink.put(MAG_VS, (BX * s4) * OBETA);
        sink.put(MAG_WA, (BX * s3) * OBETA);
        sink.put(MAG_TA, ((RI * s4) * s4) * KJ);
        sink.put(MAG_TA, ((RI * s2) * s2) * KJ);
        sink.put(MAG_TA, ((RI * s3) * s3) * KJ);
        sink.put(MAG_TA, ((RI * s2) * s4) * KJ2);
        // :609-612
        B1 -= U0 * s1;
        B1 -= V0 * s4;
        B2 -= U0 * s2;
        B3 -= U0 * s3;
        const T dxBZ = s3;
        // :614-624
        s1 = op.d2x(a.bx(1, 0, 0), BXs, a.bx(-1, 0, 0));
        s2 = op.d2x(a.by(1, 0, 0), BYs, a.by(-1, 0, 0));
        s4 = op.d2y(a.bx(0, 1, 0), BXs, a.bx(0, -1, 0));
        B1 += s1 * ORM;
        B1 += s4 * ORM;
        B2 += s2 * ORM;
        s4 = op.d2x(a.bz(1, 0, 0), BZs, a.bz(-1, 0, 0));
        B3 += s4 * ORM;
        // :626-636
        s1 = op.d1z(a.bx(0, 0, 1), a.bx(0, 0, -1));
        if constexpr (DEV) s1 += a.dzb(0);
        s2 = op.d1y(a.by(0, 1, 0), a.by(0, -1, 0));
        sink.put(MAG_UA, (BZ * s1) * OBETA);
        sink.put(MAG_WS, (BX * s1) * OBETA);
        sink.put(MAG_TA, ((RI * s1) * s1) * KJ);
        sink.put(MAG_TA, ((RI * s1) * dxBZ) * KJ2);
        B1 -= W0 * s1;
        B2 -= V0 * s2;
        // :638-658, one-sided D2z(BX) at the walls
        if (WALL == 1) {
            const T b0 = BXs, b1 = a.bx(0, 0, 1), b2 = a.bx(0, 0, 2), b3 = a.bx(0, 0, 3);
            s1 = (T(-3) * b0 + T(4) * b1 - b2) * T(c.hz) * op.d2zz +
                 (T(2) * b0 - T(5) * b1 + T(4) * b2 - b3) * T(c.h2z) * op.dzz * op.dzz;
        } else if (WALL == 2) {
            const T b0 = BXs, b1 = a.bx(0, 0, -1), b2 = a.bx(0, 0, -2), b3 = a.bx(0, 0, -3);
            s1 = (T(3) * b0 - T(4) * b1 + b2) * T(c.hz) * op.d2zz +
                 (T(2) * b0 - T(5) * b1 + T(4) * b2 - b3) * T(c.h2z) * op.dzz * op.dzz;
        } else {
            s1 = op.d2z(a.bx(0, 0, 1), BXs, a.bx(0, 0, -1));
        }
        if constexpr (DEV) s1 += a.d2zb(0);
        s2 = op.d2y(a.by(0, 1, 0), BYs, a.by(0, -1, 0));
        B1 += s1 * ORM;
        B2 += s2 * ORM;
        // :659-671
        s2 = op.d1z(a.by(0, 0, 1), a.by(0, 0, -1));
        if constexpr (DEV) s2 += a.dzb(1);
        s3 = op.d1y(a.bz(0, 1, 0), a.bz(0, -1, 0));
        s1 = s2;
        s1 -= s3;
        sink.put(MAG_VA, (BZ * s1) * OBETA);
        sink.put(MAG_WS, (BY * s1) * OBETA);
        sink.put(MAG_TA, ((RI * s2) * s2) * KJ);
        sink.put(MAG_TA, ((RI * s3) * s3) * KJ);
        sink.put(MAG_TA, ((RI * s2) * s3) * KJ2);
        B2 -= W0 * s2;
        B3 -= V0 * s3;
        // :672-690, one-sided D2z(BY) at the walls
        if (WALL == 1) {
            const T b0 = BYs, b1 = a.by(0, 0, 1), b2 = a.by(0, 0, 2), b3 = a.by(0, 0, 3);
            s2 = (T(-3) * b0 + T(4) * b1 - b2) * T(c.hz) * op.d2zz +
                 (T(2) * b0 - T(5) * b1 + T(4) * b2 - b3) * T(c.h2z) * op.dzz * op.dzz;
        } else if (WALL == 2) {
            const T b0 = BYs, b1 = a.by(0, 0, -1), b2 = a.by(0, 0, -2), b3 = a.by(0, 0, -3);
            s2 = (T(3) * b0 - T(4) * b1 + b2) * T(c.hz) * op.d2zz +
                 (T(2) * b0 - T(5) * b1 + T(4) * b2 - b3) * T(c.h2z) * op.dzz * op.dzz;
        } else {
            s2 = op.d2z(a.by(0, 0, 1), BYs, a.by(0, 0, -1));
        }
        if constexpr (DEV) s2 += a.d2zb(1);
        s3 = op.d2y(a.bz(0, 1, 0), BZs, a.bz(0, -1, 0));
        B2 += s2 * ORM;
        B3 += s3 * ORM;
        // :691-694
        s3 = op.d1z(a.bz(0, 0, 1), a.bz(0, 0, -1));
        if constexpr (DEV) s3 += a.dzb(2);
        B3 -= W0 * s3;
        s3 = op.d2z(a.bz(0, 0, 1), BZs, a.bz(0, 0, -1));
        if constexpr (DEV) s3 += a.d2zb(2);
        B3 += s3 * ORM;
    }
}

// fluxes() of one point: the hydrodynamic part, then (lMag) the magnetic
// part with its contributions applied in order.
template <bool MAG, bool ROT, int WALL, typename T, bool DEV = false, bool EXT = false,
          int EXTM = EXT_INTERIOR, typename A>
__device__ __forceinline__ void flux_point(const A& a, const Phys& c, const PointMetric& m,
                                           FluxOutT<T>& o, const ExtPt* e = nullptr)
{
    T FU, FV, FW, FR, FT;
    flux_hydro<ROT, WALL, T, DEV, A, EXT, EXTM>(a, c, m, FU, FV, FW, FR, FT, e);
    T B1 = T(0), B2 = T(0), B3 = T(0);
    if (MAG) {
        InlineSink<T> sink{FU, FV, FW, FT};
        flux_mag<WALL, T, DEV>(a, c, m, sink, B1, B2, B3);
    }
    o.FU = FU;
    o.FV = FV;
    o.FW = FW;
    o.FR = FR;
    o.FT = FT;
    o.B1 = B1;
    o.B2 = B2;
    o.B3 = B3;
}

}  // namespace mhd
