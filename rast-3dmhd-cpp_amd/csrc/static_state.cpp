// The reference's initial state Static() (3dmhdphsic.h:549-763), on the host:
// the polytrope of the z grid's hydrostatic ODE (derivs, :350-358) integrated
// by Bulirsch-Stoer steps (bsstep/mmid/rzextr, :229-348) -- or, with lRem,
// M. Rempel's background (getbackground/ODEINT/RKQC/RK4/derivs1/kappa,
// :357-522) with its conductivity profile and wall temperatures, or with lShr
// a uniform state -- each rank's planes taken as that rank takes them, the
// magnetic layer of lMag (the AMPB Gaussian, :682-699, or tube() with
// nTube = 0, :763-781), and the RAN2 perturbation (:523-547) drawn in the rank
// loop order of rank 0 (:705-738), applied to TT (to RV with lShr).  The result
// is the state the reference's nPEy x nPEz run starts from, assembled into one
// domain, so that run can be continued on one GPU.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mhd3d.h"

namespace {

// RAN2 (3dmhdphsic.h:523-547): shuffle table of a linear congruential
// generator; the table and IIY persist between calls like the reference's
// statics.
class Ran2 {
public:
    explicit Ran2(int idum) : idum_(idum) {}
    bool next(double& out)
    {
        const int M = 714025, IA = 1366, IC = 150889;
        const double RM = (double)(1.0f / (float)M);
        if (idum_ < 0) {
            idum_ = (int)std::fmod((double)(IC - idum_), (double)M);
            for (int j = 0; j < 97; ++j) {
                idum_ = (int)std::fmod((double)(IA * idum_ + IC), (double)M);
                table_[j] = idum_;
            }
            idum_ = (int)std::fmod((double)(IA * idum_ + IC), (double)M);
            iiy_ = idum_;
        }
        const int j = (97 * iiy_) / M;
        if (j >= 97 || j < 0) return false;   // "RAN2: Error." in the reference
        iiy_ = table_[j];
        out = (double)(float)iiy_ * RM;
        idum_ = (IA * idum_ + IC) % M;
        table_[j] = idum_;
        return true;
    }

private:
    int idum_;
    int iiy_ = 0;
    int table_[97] = {0};
};

// Bulirsch-Stoer integration of derivs with nGrid = 0 and PZP = 0 (the
// stratification without a stable layer): dT/dz = THETA/dzzdz,
// d ln(RO*T)/dz = GRAV/T/dzzdz with dzzdz = 1/zMax.
class Polytrope {
public:
    Polytrope(double theta, double grav, double zmax) : theta_(theta), grav_(grav), zmax_(zmax) {}

    void derivs(const double* y, double* dydx) const
    {
        const double dzzdz = 1.0 / zmax_;
        const double rkap = 1.0;
        dydx[0] = theta_ / rkap / dzzdz;
        dydx[1] = grav_ / (1.0 + y[0]) / dzzdz;
    }

    // Modified midpoint over htot in nstp substeps (mmid).
    void midpoint(const double* y, const double* dydx, double htot, int nstp, double* yout)
    {
        const double h = htot / nstp;
        for (int i = 0; i < 2; ++i) {
            ym_[i] = y[i];
            yn_[i] = y[i] + h * dydx[i];
        }
        derivs(yn_, yout);
        const double h2 = 2.0 * h;
        for (int n = 2; n <= nstp; ++n) {
            for (int i = 0; i < 2; ++i) {
                const double swap = ym_[i] + h2 * yout[i];
                ym_[i] = yn_[i];
                yn_[i] = swap;
            }
            derivs(yn_, yout);
        }
        for (int i = 0; i < 2; ++i) yout[i] = 0.5 * (ym_[i] + yn_[i] + h * yout[i]);
    }

    // Rational extrapolation to h -> 0 (rzextr); x_ and fx_ persist like the
    // reference's statics.
    void extrapolate(int iest, double xest, const double* yest, double* yz, double d[][7],
                     double* dy, int nuse)
    {
        x_[iest] = xest;
        if (iest == 0) {
            for (int j = 0; j < 2; ++j) yz[j] = d[j][0] = dy[j] = yest[j];
            return;
        }
        const int m1 = iest + 1 < nuse ? iest + 1 : nuse;
        for (int k = 1; k < m1; ++k) fx_[k] = x_[iest - k] / xest;
        for (int j = 0; j < 2; ++j) {
            double yy = yest[j], v = d[j][0], c = yy, ddy = 0.0;
            d[j][0] = yy;
            for (int k = 1; k < m1; ++k) {
                const double b1 = fx_[k] * v;
                double b = b1 - c;
                if (b != 0.0) {
                    b = (c - v) / b;
                    ddy = c * b;
                    c = b1 * b;
                } else {
                    ddy = v;
                }
                v = d[j][k];
                d[j][k] = ddy;
                yy = yy + ddy;
            }
            dy[j] = ddy;
            yz[j] = yy;
        }
    }

    // One bsstep of size htry; false if the step underflows.
    bool step(double* y, const double* dydx, double& x, double htry, double eps, double& hdid)
    {
        static const int kSeq[11] = {2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96};
        const int imax = 11, nuse = 7;
        double h = htry;
        const double xsav = x;
        const double ysav[2] = {y[0], y[1]}, dysav[2] = {dydx[0], dydx[1]};
        double d[2][7];
        std::memset(d, 0, sizeof(d));
        for (int tries = 0; tries < 64; ++tries) {
            for (int i = 0; i < imax; ++i) {
                double yseq[2], yerr[2];
                midpoint(ysav, dysav, h, kSeq[i], yseq);
                const double q = h / kSeq[i];
                extrapolate(i, q * q, yseq, y, d, yerr, nuse);
                double errmax = 0.0;
                for (int j = 0; j < 2; ++j) {
                    const double e = std::fabs(yerr[j] / y[j]);
                    if (errmax < e) errmax = e;   // std::max
                }
                errmax /= eps;
                if (errmax < 1) {
                    x = x + h;
                    hdid = h;
                    return true;
                }
            }
            h = 0.25 * h / std::pow(2.0, (imax - nuse) / 2);
            if (x + h == x) return false;
        }
        (void)xsav;
        return false;
    }

private:
    double theta_, grav_, zmax_;
    double ym_[2] = {0, 0}, yn_[2] = {0, 0};
    double x_[11] = {0}, fx_[7] = {0};
};

// M. Rempel's background (lRem): kappa (3dmhdphsic.h:390-392), derivs1
// (:394-406), the fixed-step RK4 (:508-518), the step-doubling RKQC
// (:470-506) and the driver ODEINT (:410-468) that getbackground (:357-362)
// runs from x = 2 down to the plane's height with Y = (T, P) = (1, 1).
class Rempel {
public:
    static double kappa(double z, double rho, double t)
    {
        (void)t;
        return rho + (1.25f - rho) * std::exp(-5.0f * z);
    }

    static void derivs(double x, const double* yt, double* dyt)
    {
        const double nad = 0.41f;
        const double os = 1.0f;
        const double krad = kappa(x, yt[1] / yt[0], yt[0]);
        const double nrad = 0.4f / krad;
        dyt[0] = nrad <= os * nad ? nrad : nad;
        dyt[1] = yt[1] / yt[0];
    }

    static void rk4(const double* y, const double* dydx, double x, double h, double* yout)
    {
        double yt[2], dyt[2], dym[2];
        const double hh = h * 0.5f, h6 = h / 6.0f, xh = x + hh;
        for (int i = 0; i < 2; i++) yt[i] = y[i] + hh * dydx[i];
        derivs(xh, yt, dyt);
        for (int i = 0; i < 2; i++) yt[i] = y[i] + hh * dyt[i];
        derivs(xh, yt, dym);
        for (int i = 0; i < 2; i++) yt[i] = y[i] + h * dym[i], dym[i] = dyt[i] + dym[i];
        derivs(x + h, yt, dyt);
        for (int i = 0; i < 2; i++) yout[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0f * dym[i]);
    }

    // one accepted step of at most htry; false: "Stepsize not significant"
    static bool rkqc(double* y, double* dydx, double& x, double htry, double eps, const double* yscal,
                     double& hdid, double& hnext)
    {
        const double FCOR = 1.0f / 15.0f;
        const double SAFETY = 0.9f;
        const double ERRCON = 6.0e-4f;
        const double PGROW = -0.20f;
        const double PSHRNK = -0.25f;
        const double xsav = x;
        double ysav[2] = {y[0], y[1]}, dysav[2] = {dydx[0], dydx[1]}, ytemp[2];
        double h = htry;
        for (;;) {
            const double hh = 0.5f * h;
            rk4(ysav, dysav, xsav, hh, ytemp);
            x = xsav + hh;
            derivs(x, ytemp, dydx);
            rk4(ytemp, dydx, x, hh, y);
            x = xsav + h;
            if (x == xsav) return false;
            rk4(ysav, dysav, xsav, h, ytemp);
            double errmax = 0.0f;
            for (int i = 0; i < 2; i++) {
                ytemp[i] = y[i] - ytemp[i];
                const double e = std::fabs(ytemp[i] / yscal[i]);
                if (errmax < e) errmax = e;   // std::max
            }
            errmax = errmax / eps;
            if (errmax > 1.0) {
                h = SAFETY * h * std::pow(errmax, PSHRNK);
            } else {
                hdid = h;
                hnext = errmax > ERRCON ? SAFETY * h * std::pow(errmax, PGROW) : 4.0f * h;
                break;
            }
        }
        for (int i = 0; i < 2; i++) y[i] += ytemp[i] * FCOR;
        return true;
    }

    // getbackground(x): T and rho at height x; false on an integrator failure
    static bool background(double x2, double& t, double& rho)
    {
        const int MAXSTP = 1000;
        const double TINY = 1.0e-30f, EPS = 1e-8, H1 = 1e-4, HMIN = 1e-50;
        const double x1 = 2.0f;
        double y[2] = {1.0, 1.0}, dydx[2], yscal[2];
        double x = x1, h = (x2 > x1) ? H1 : -H1;
        for (int n = 1; n <= MAXSTP; ++n) {
            derivs(x, y, dydx);
            for (int i = 0; i < 2; ++i) yscal[i] = std::fabs(y[i]) + std::fabs(h * dydx[i]) + TINY;
            if ((x + h - x2) * (x + h - x1) > 0) h = x2 - x;
            double hdid = 0.0, hnext = 0.0;
            if (!rkqc(y, dydx, x, h, EPS, yscal, hdid, hnext)) return false;
            if ((x - x2) * (x2 - x1) >= 0) {
                t = y[0];
                rho = y[1] / y[0];
                return true;
            }
            if (std::fabs(hnext) < HMIN) return false;
            h = hnext;
        }
        return false;
    }
};

}  // namespace

extern "C" int mhd3d_static_state(int nPx, int nPy, int nPz_global, int nPEy, int nPEz,
                                  const double* par, int lMag, double* const* fields)
{
    return mhd3d_static_state_flags(nPx, nPy, nPz_global, nPEy, nPEz, par, lMag, 0, 0, fields,
                                    nullptr, nullptr, nullptr);
}

extern "C" int mhd3d_static_state_flags(int nPx, int nPy, int nPz_global, int nPEy, int nPEz,
                                        const double* par, int lMag, int lShr, int lRem,
                                        double* const* fields, double* RKAPA, double* DKAPA,
                                        double* walls)
{
    if (nPx < 1 || nPy < 1 || nPz_global < 3 || nPEy < 1 || nPEz < 1 || !par || !fields)
        return MHD3D_EINVAL;
    if (lShr && lRem) return MHD3D_EUNSUPPORTED;   // refused by the timestep as well
    if (lRem && (!RKAPA || !DKAPA || !walls)) return MHD3D_EINVAL;
    if (nPy % nPEy != 0) return MHD3D_EINVAL;
    const int nf = lMag ? 8 : 5;
    for (int f = 0; f < nf; ++f)
        if (!fields[f]) return MHD3D_EINVAL;
    const int nRz = nPz_global / nPEz, nRy = nPy / nPEy;
    const int nzd = nPEz * nRz;   // planes of the assembled domain (incl. both walls)
    if (nRz < 1) return MHD3D_EINVAL;
    const double THETA = par[2], GRAV = par[3], zMax = par[18], AMPT = par[19];

    // the background on the global z grid (:565-616)
    std::vector<double> T(nPz_global), R(nPz_global), RK(nPz_global), DRK(nPz_global);
    if (lShr) {
        for (int k = 0; k < nPz_global; ++k) T[k] = 1.0f, R[k] = 1.0f;
    } else if (lRem) {
        // mkGrid (nGrid 0, :37-41): height s*zMax, ds/dz = 1.0f/zMax
        const double REPR = par[0] * par[1];                  // 3dmhd.cpp:451
        const double hz = (double)(0.5f * (float)(nPz_global - 1));   // 3dmhd.cpp:368
        const double sdz = 1.0f / zMax;
        for (int k = 0; k < nPz_global; ++k) {
            const double sz = (double)((float)k / (float)(nPz_global - 1)) * zMax;
            if (!Rempel::background(sz, T[k], R[k])) return MHD3D_EINVAL;
            RK[k] = Rempel::kappa(sz, R[k], T[k]);
            RK[k] *= 8.07f * THETA * REPR;
        }
        DRK[0] = (-3.0f * RK[0] + 4.0f * RK[1] - RK[2]) * hz * sdz;
        for (int k = 1; k < nPz_global - 1; ++k) DRK[k] = (RK[k + 1] - RK[k - 1]) * hz * sdz;
        const int n = nPz_global;
        DRK[n - 1] = (3.0f * RK[n - 1] - 4.0f * RK[n - 2] + RK[n - 3]) * hz * sdz;
        walls[0] = T[0];                                                      // TU
        walls[1] = T[0] - 4.0f / 3.0f * T[1] + 1.0f / 3.0f * T[2];           // DZTU
        walls[2] = T[n - 1] - 4.0f / 3.0f * T[n - 2] + 1.0f / 3.0f * T[n - 3];   // DZTB
    } else {
        // the polytrope (:595-609)
        Polytrope poly(THETA, GRAV, zMax);
        const double eps = (double)1.0E-12f;
        const double dzz = (double)(1.0f / (float)(nPz_global - 1));
        double y[2] = {0.0, 0.0}, dydx[2];
        T[0] = 1.0;
        R[0] = 1.0;
        double zz = 0.0;
        for (int k = 1; k < nPz_global; ++k) {
            double hdid = 0.0;
            poly.derivs(y, dydx);
            if (!poly.step(y, dydx, zz, dzz, eps, hdid) || hdid != dzz) return MHD3D_EINVAL;   // "Static structure error"
            T[k] = 1.0 + y[0];
            R[k] = std::exp(y[1]) / (1.0 + y[0]);
        }
    }

    // each plane as its rank takes it (:613-680): ranks below the last take
    // T[g] for global interior plane g; the last rank aligns its planes with
    // the top of T (so a remainder nPz_global - nPEz*nRz is skipped there)
    const size_t nx = (size_t)nPx + 2, ny = (size_t)nPy + 2, nz = (size_t)nzd + 2;
    const size_t arr = nx * ny * nz;
    for (int f = 0; f < nf; ++f) std::memset(fields[f], 0, arr * sizeof(double));
    double* RO = fields[MHD3D_RO];
    double* TT = fields[MHD3D_TT];
    // global background index of each plane of the assembled domain
    std::vector<int> kmap(nz);
    for (size_t z = 0; z < nz; ++z) {
        int k;
        if (z == 0) {
            k = 0;
        } else if (z == nz - 1) {
            k = nPz_global - 1;
        } else {
            const int g = (int)z - 1;
            k = g / nRz < nPEz - 1 ? g : g + (nPz_global - nzd);
        }
        kmap[z] = k;
        for (size_t yy = 1; yy <= (size_t)nPy; ++yy)
            for (size_t xx = 1; xx <= (size_t)nPx; ++xx) {
                TT[(z * ny + yy) * nx + xx] = T[k];
                RO[(z * ny + yy) * nx + xx] = R[k];
            }
        if (lRem) {
            RKAPA[z] = RK[k];
            DKAPA[z] = DRK[k];
        }
    }

    // the magnetic layer (:682-703): every rank sets BX on its local planes
    // 2 .. nRz+1 (its first interior plane stays 0) from its own ZEE, and
    // lowers RO there by BX^2/TT*OBETA; nTube = 0 leaves B = 0 otherwise
    const double AMPB = par[20], BFH = par[21], BZP = par[22];
    if (lMag && AMPB != 0.0) {
        const double OBETA = 2.0f / par[7];                   // 3dmhd.cpp:459
        const double ORZ = 1.0f / (float)(nPz_global - 1);   // zJacobi (:125-156)
        const double CLN = -4.0f * std::log(2.0f) / (BFH * BFH);
        double* BX = fields[5];
        for (size_t z = 1; z < nz; ++z) {
            // the plane's owner and local index; the bottom ghost is the last
            // rank's local plane nRz+1
            const int g = (int)z - 1;
            const int local = z == nz - 1 ? nRz + 1 : g % nRz + 1;
            if (local < 2) continue;
            const double zee = z == nz - 1 ? 1.0f * zMax : ((double)(float)kmap[z] * ORZ) * zMax;
            const double b = AMPB * std::exp(CLN * ((zee - BZP) * (zee - BZP)));
            for (size_t yy = 0; yy < ny; ++yy)
                for (size_t xx = 0; xx < nx; ++xx) BX[(z * ny + yy) * nx + xx] = b;
            for (size_t yy = 1; yy <= (size_t)nPy; ++yy)
                for (size_t xx = 1; xx <= (size_t)nPx; ++xx) {
                    const size_t i = (z * ny + yy) * nx + xx;
                    RO[i] -= (BX[i] * BX[i] / TT[i]) * OBETA;
                }
        }
    }

    // RAN2 perturbation of TT (:705-738): rank 0 draws, for every z rank,
    // local plane, y rank, row and column in that order
    if (AMPT != 0.0) {
        Ran2 rng(-62659);
        for (int pz = 0; pz < nPEz; ++pz)
            for (int k = 1; k <= nRz; ++k)
                for (int py = 0; py < nPEy; ++py)
                    for (int yy = 1; yy <= nRy; ++yy)
                        for (int xx = 1; xx <= nPx; ++xx) {
                            double r;
                            if (!rng.next(r)) return MHD3D_EINVAL;
                            const size_t gz = (size_t)pz * nRz + k;
                            const size_t gy = (size_t)py * nRy + yy;
                            const size_t i = (gz * ny + gy) * nx + xx;
                            const double w = 1.0 + AMPT * (r - 0.5E00);
                            if (lShr)
                                fields[MHD3D_RV][i] = fields[MHD3D_RV][i] + w - 1.0f;
                            else
                                TT[i] = TT[i] * w;
                        }
    }
    // x and y ghosts periodic (communicate(), 3dmhdmpi.h:59-142)
    for (int f = 0; f < nf; ++f) {
        double* a = fields[f];
        for (size_t z = 0; z < nz; ++z) {
            double* pl = a + z * ny * nx;
            for (size_t yy = 1; yy <= (size_t)nPy; ++yy) {
                pl[yy * nx] = pl[yy * nx + nx - 2];
                pl[yy * nx + nx - 1] = pl[yy * nx + 1];
            }
            std::memcpy(pl, pl + (ny - 2) * nx, nx * sizeof(double));
            std::memcpy(pl + (ny - 1) * nx, pl + nx, nx * sizeof(double));
        }
    }
    return MHD3D_OK;
}
