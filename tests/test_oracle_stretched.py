"""The oracle on a stretched z grid and the magnetic CFL limit (CPU).

Every workload of the benchmark runs on a tanh-stretched z grid, where the
metric terms of the z operators matter: D1z f = (f[k+1]-f[k-1]) dzzdz hz and
D2z f = (f[k+1]-f[k-1]) d2zzdz2 hz + (f[k+1]-2f[k]+f[k-1]) dzzdz^2 h2z
(3dmhdutils.h:195-239), and the one-sided wall stencils carry the same two
metric factors (3dmhdsub.h:211-226,572-584,640-654,673-687).  On a uniform
grid d2zzdz2 = 0 and dzzdz is constant, so none of that is seen there.

What fixes the answer here, independently of the oracle:

* a z map z = Z(s) of the computational coordinate s written in sympy; its
  inverse derivatives ds/dz = 1/Z'(s) and d2s/dz2 = (ds/dz) d/ds(ds/dz) are
  taken symbolically;
* fields that are polynomials of low degree in (x, y, s): then the central
  differences in s are exact, and the chain rule makes every discrete z
  derivative equal the physical one -- so every flux of the oracle, FR, FU,
  FV, FW, FT and the induction WW1-3 with rotation and the magnetic block,
  must equal the continuum expression of the same equations written in
  sympy (physical derivatives d/dz = (ds/dz) d/ds), to rounding, inside the
  domain and on both walls;
* the product's own tanh grid (``tanh_axis``) agrees with the symbolic map,
  and on it D1z and D2z of smooth non-polynomial fields converge to the
  analytic derivatives at second order (first order for the one-sided D2z
  on the walls);
* the CFL limits (3dmhdsub.h:70-105): each of the four minima equals an
  independent numpy minimum over the grid, and on a state where the magnetic
  limit 0.5 DD^2 RM binds, DT = SFF * 0.5 * DDmin^2 * RM.
"""
import numpy as np
import pytest
import sympy as sp

from mhd3d.desc import Descriptor, RunConfig, default_par, fdiv, tanh_axis
from oracle import oracle as om

NPX, NPY, NPZ = 12, 10, 11
BETA = 1.5          # the benchmark's stretch factor
ZMAX = 1.0

X, Y, S = sp.symbols("x y s", real=True)


def _zmap(zmax=ZMAX, beta=BETA):
    """z(s) of the symmetric tanh stretching, clustered at both walls."""
    return zmax / 2 * (1 + sp.tanh(beta * (2 * S - 1)) / sp.tanh(beta))


def _inverse_derivatives(Z):
    """ds/dz and d2s/dz2 as functions of s, from the map z = Z(s)."""
    s1 = 1 / sp.diff(Z, S)
    s2 = s1 * sp.diff(s1, S)
    return s1, s2


def _par(RM=25.0):
    par = default_par()
    par[0] = 40.0          # RE
    par[1] = 0.7           # PR
    par[2] = 3.0           # THETA
    par[3] = 8.0           # GRAV
    par[4] = 2.0           # RY
    par[5] = 0.3           # ANG
    par[6] = RM            # RM
    par[7] = 1.5           # BETA
    par[16], par[17], par[18] = 2.0, 1.5, ZMAX
    return par


def _mapped_config(lMag=True, lRot=True, RM=25.0):
    """A run whose z metric arrays are those of the symbolic map, sampled at
    the stencil's own computational points s_k = (k-1)/(nPz-1) (ghosts
    included), and whose conductivity profile varies with z."""
    Z = _zmap()
    s1, s2 = _inverse_derivatives(Z)
    sk = (np.arange(NPZ + 2) - 1.0) / (NPZ - 1)
    ZEE = sp.lambdify(S, Z, "numpy")(sk)
    dzzdz = sp.lambdify(S, s1, "numpy")(sk)
    d2zzdz2 = sp.lambdify(S, s2, "numpy")(sk)
    OR = fdiv(1.0, float(NPZ - 1))
    DDZ = OR / dzzdz
    rk = 1.0 + 0.4 * sk - 0.2 * sk ** 2
    dk = 0.3 - 0.5 * sk
    cfg = RunConfig(nPx=NPX, nPy=NPY, nPz=NPZ, lMag=lMag, lRot=lRot, izCon=1, iTCon=4,
                    par=_par(RM), zgrid=("arrays", ZEE, dzzdz, d2zzdz2, DDZ), kappa=(rk, dk))
    return cfg, sk, (Z, s1, s2), (rk, dk)


def _grid(desc, sk):
    """Grid points the stencils imply: uniform x, y spacing 1/(2 h dxxdx), and
    the computational coordinate s_k along z."""
    c, m = desc.consts, desc.metrics
    dx = 1.0 / (2.0 * c["hx"] * m.dxxdx[1])
    dy = 1.0 / (2.0 * c["hy"] * m.dyydy[1])
    nz, ny, nx = desc.shape
    Sg, Yg, Xg = np.meshgrid(sk, (np.arange(ny) - 1.0) * dy, (np.arange(nx) - 1.0) * dx,
                             indexing="ij")
    return Xg, Yg, Sg


def _rand_poly(rng, deg, scale):
    """A random polynomial with degree at most deg[i] in each of x, y, s."""
    e = 0
    for i in range(deg[0] + 1):
        for j in range(deg[1] + 1):
            for k in range(deg[2] + 1):
                e += sp.Float(rng.uniform(-1, 1) * scale) * X ** i * Y ** j * S ** k
    return e


def _fields(seed, wall=None):
    """Velocity multilinear in (x, y, s) (its products with itself stay
    quadratic in each variable), temperature and field quadratic in each.
    With wall = 0 or 1, w vanishes on that wall (s = wall), as RW = 0 there
    after BCON -- the one-sided stencils assume it."""
    rng = np.random.default_rng(seed)
    u = 0.05 * _rand_poly(rng, (1, 1, 1), 1.0)
    v = 0.05 * _rand_poly(rng, (1, 1, 1), 1.0)
    if wall is None:
        w = 0.05 * _rand_poly(rng, (1, 1, 1), 1.0)
    else:
        w = 0.05 * (S - wall) * _rand_poly(rng, (1, 1, 0), 1.0)
    T = 2.0 + 0.3 * _rand_poly(rng, (2, 2, 2), 0.5)
    B = [0.02 * _rand_poly(rng, (2, 2, 2), 0.5) for _ in range(3)]
    return u, v, w, T, B


def _continuum(desc, s1, rk, dk, rho, u, v, w, T, B, mag, rot):
    """The flux of every variable from the continuum equations the reference
    discretises (3dmhdsub.h:204-695), with d/dz = (ds/dz) d/ds."""
    c = desc.consts
    kap = sp.Function("kappa")(S)
    dkap = sp.Function("dkappa")(S)

    def dx(f):
        return sp.diff(f, X)

    def dy(f):
        return sp.diff(f, Y)

    def dz(f):
        return s1 * sp.diff(f, S)

    def lap(f):
        return dx(dx(f)) + dy(dy(f)) + dz(dz(f))

    U = [u, v, w]
    D = [dx, dy, dz]
    RU = [rho * q for q in U]
    ORE, C43, C13, OCV = c["ORE"], c["C43"], c["C13"], c["OCV"]
    div = dx(u) + dy(v) + dz(w)
    P = rho * T
    F = {"FR": -(dx(RU[0]) + dy(RU[1]) + dz(RU[2]))}
    mom = []
    for i in range(3):
        adv = sum(D[j](RU[i] * U[j]) for j in range(3))
        visc = ORE * (lap(U[i]) + (C43 - 1.0) * D[i](D[i](U[i])))
        cross = C13 * ORE * sum(D[i](D[j](U[j])) for j in range(3) if j != i)
        mom.append(-D[i](P) - adv + visc + cross)
    mom[2] += c["GRAV"] * rho
    if rot:
        mom[0] += c["OMZ"] * RU[1]
        mom[1] += -c["OMZ"] * RU[0] + c["OMX"] * RU[2]
        mom[2] += -c["OMX"] * RU[1]
    g = [[D[j](U[i]) for j in range(3)] for i in range(3)]    # g[i][j] = d u_i / d x_j
    phi = (C43 * (g[0][0] ** 2 + g[1][1] ** 2 + g[2][2] ** 2
                  - g[0][0] * g[1][1] - g[1][1] * g[2][2] - g[2][2] * g[0][0])
           + (g[0][1] + g[1][0]) ** 2 + (g[0][2] + g[2][0]) ** 2 + (g[1][2] + g[2][1]) ** 2)
    FT = (-(u * dx(T) + v * dy(T) + w * dz(T))
          + OCV / c["REPR"] / rho * (kap * (dx(dx(T)) + dy(dy(T)) + dz(dz(T))) + dkap * dz(T))
          + OCV * ORE * phi / rho - OCV * T * div)
    if mag:
        J = [dy(B[2]) - dz(B[1]), dz(B[0]) - dx(B[2]), dx(B[1]) - dy(B[0])]
        lor = [J[1] * B[2] - J[2] * B[1], J[2] * B[0] - J[0] * B[2], J[0] * B[1] - J[1] * B[0]]
        for i in range(3):
            mom[i] += c["OBETA"] * lor[i]
        FT += c["OBETA"] * c["ORM"] * OCV * (J[0] ** 2 + J[1] ** 2 + J[2] ** 2) / rho
        for i in range(3):
            F[f"WW{i + 1}"] = (sum(B[j] * g[i][j] for j in range(3)) - B[i] * div
                               - sum(U[j] * D[j](B[i]) for j in range(3)) + c["ORM"] * lap(B[i]))
    F["FU"], F["FV"], F["FW"] = mom
    F["FT"] = FT
    out = {}
    for k, e in F.items():
        out[k] = e.subs({kap: sp.Symbol("KAP"), dkap: sp.Symbol("DKAP")})
    return out


def _evaluate(expr, Xg, Yg, Sg, rk, dk):
    f = sp.lambdify((X, Y, S, sp.Symbol("KAP"), sp.Symbol("DKAP")), expr, "numpy")
    K = np.broadcast_to(rk[:, None, None], Xg.shape)
    DK = np.broadcast_to(dk[:, None, None], Xg.shape)
    return np.broadcast_to(f(Xg, Yg, Sg, K, DK), Xg.shape)


def _run_case(seed, wall=None, mag=True, rot=True, rho_z=False):
    cfg, sk, (Z, s1, s2), (rk, dk) = _mapped_config(lMag=mag, lRot=rot)
    desc = Descriptor(cfg)
    Xg, Yg, Sg = _grid(desc, sk)
    u, v, w, T, B = _fields(seed, wall)
    rho = sp.Float(1.3)
    if rho_z:
        # density varying along z with the velocity independent of s keeps
        # every differenced product quadratic in s
        rho = 1.3 + 0.4 * S
        u, v = u.subs(S, 0), v.subs(S, 0)
        w = sp.Float(0.0)
        T = 2.0 + 0.5 * S + 0.2 * X - 0.1 * Y
    exprs = {"RO": rho, "TT": T, "RU": rho * u, "RV": rho * v, "RW": rho * w}
    if mag:
        exprs.update({"BX": B[0], "BY": B[1], "BZ": B[2]})
    st = {k: np.array(_evaluate(e, Xg, Yg, Sg, rk, dk)) for k, e in exprs.items()}
    o = om.Oracle(desc)
    o.set_state(st)                 # no periodic refill: polynomial ghosts
    o.fluxes()
    E = _continuum(desc, s1, rk, dk, rho, u, v, w, T, B, mag, rot)
    names = ["FR", "FU", "FV", "FW", "FT"] + (["WW1", "WW2", "WW3"] if mag else [])
    got = {n: o.get(n) for n in names}
    want = {n: _evaluate(E[n], Xg, Yg, Sg, rk, dk) for n in names}
    o.close()
    return got, want


def _check(got, want, sl, names, tol=1e-10):
    for n in names:
        g, e = got[n][sl], want[n][sl]
        err = np.abs(g - e).max() / max(np.abs(e).max(), 1e-30)
        assert err < tol, (n, err)


@pytest.mark.parametrize("seed", [1, 2])
def test_every_flux_equals_its_continuum_value_on_a_mapped_grid(seed):
    got, want = _run_case(seed)
    # planes between the walls, points away from the (non-periodic) ghosts
    sl = (slice(2, -2), slice(2, -2), slice(2, -2))
    _check(got, want, sl, ["FR", "FU", "FV", "FW", "FT", "WW1", "WW2", "WW3"])


def test_density_varying_along_the_map():
    """RO(s) enters the pressure, the advection products and every 1/RO
    factor (conduction, dissipation) on the mapped grid."""
    got, want = _run_case(5, mag=False, rot=False, rho_z=True)
    sl = (slice(2, -2), slice(2, -2), slice(2, -2))
    _check(got, want, sl, ["FR", "FU", "FV", "FW", "FT"])


@pytest.mark.parametrize("wall", [0, 1])
def test_wall_planes_on_a_mapped_grid(wall):
    """Top (s = 0, plane 1) and bottom (s = 1, plane nPz) walls: the
    one-sided D1z of RW in the continuity equation and of w in the induction,
    and the one-sided D2z of BX and BY, with both metric factors."""
    got, want = _run_case(11 + wall, wall=wall)
    z = 1 if wall == 0 else NPZ
    sl = (z, slice(2, -2), slice(2, -2))
    _check(got, want, sl, ["FR", "FU", "FV", "FT", "WW1", "WW2", "WW3"])


def test_tanh_axis_is_the_symbolic_map():
    """The product's z grid: ZEE = Z(s), dzzdz = 1/Z'(s), d2zzdz2 = the
    second derivative of the inverse map, at s = k * OR (OR = 1.0f/(nPz-1)
    in float, as the reference's Jacobi routines form it), ghosts clamped to
    the walls; DDZ = OR / dzzdz (3dmhdphsic.h:155)."""
    for n in (9, 64, 128):
        Z = _zmap(2.5, BETA)
        s1, s2 = _inverse_derivatives(Z)
        ZEE, du, d2, dd = tanh_axis(n, 2.5, BETA)
        OR = fdiv(1.0, float(n - 1))
        k = np.clip(np.arange(n + 2) - 1, 0, n - 1)
        s = k * OR
        for got, e in ((ZEE, Z), (du, s1), (d2, s2)):
            want = sp.lambdify(S, e, "numpy")(s)
            assert np.allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
        assert np.array_equal(dd, OR * (1.0 / du))
        assert ZEE[1] == pytest.approx(0.0, abs=1e-12)
        assert ZEE[n] == pytest.approx(2.5, rel=1e-6)
        assert np.all(np.diff(ZEE[1:n + 1]) > 0)


def _tanh_desc(nPz, lMag=False, kappa=None):
    par = _par()
    cfg = RunConfig(nPx=8, nPy=6, nPz=nPz, lMag=lMag, izCon=1, iTCon=4, par=par,
                    zgrid=("tanh", BETA), kappa=kappa)
    return Descriptor(cfg)


def _zfield(desc, f):
    nz, ny, nx = desc.shape
    z = desc.metrics.ZEE
    return np.broadcast_to(f(z)[:, None, None], (nz, ny, nx)).copy()


def test_z_derivatives_converge_on_the_tanh_grid():
    """On the product's tanh grid, D1z (through the continuity equation and
    the DKAPA conduction term) and D2z (conduction) of smooth fields of the
    physical z converge to the analytic derivatives at second order; a
    dropped or misplaced metric term leaves an O(1) error."""
    L = ZMAX
    errs = []
    for n in (17, 33, 65):
        nz = n + 2
        kap = (np.ones(nz), np.full(nz, 0.5))
        desc = _tanh_desc(n, kappa=kap)
        c = desc.consts
        zeros = np.zeros(desc.shape)
        w = lambda z: 0.3 * np.sin(np.pi * z / L) * (1 + 0.5 * z)
        dw = lambda z: 0.3 * (np.pi / L * np.cos(np.pi * z / L) * (1 + 0.5 * z) + 0.5 * np.sin(np.pi * z / L))
        T = lambda z: 1.0 + 2.0 * z + 0.1 * np.cos(2 * np.pi * z / L)
        dT = lambda z: 2.0 - 0.2 * np.pi / L * np.sin(2 * np.pi * z / L)
        d2T = lambda z: -0.4 * (np.pi / L) ** 2 * np.cos(2 * np.pi * z / L)
        st = {"RO": np.ones(desc.shape), "TT": _zfield(desc, T), "RU": zeros, "RV": zeros,
              "RW": _zfield(desc, w)}
        o = om.Oracle(desc)
        o.set_state(st)
        o.fluxes()
        z = desc.metrics.ZEE
        FR = o.get("FR")[1:n + 1, 1:-1, 1:-1]
        # FR = -d(RW)/dz with one-sided stencils on the walls; FT = conduction
        # minus advection and pressure work of the same RW
        eFR = np.abs(FR + dw(z[1:n + 1])[:, None, None]).max()
        o.set_state({"RW": zeros})
        o.fluxes()
        FT = o.get("FT")[2:n, 1:-1, 1:-1]
        want = c["OCV"] / c["REPR"] * (d2T(z[2:n]) + 0.5 * dT(z[2:n]))
        eFT = np.abs(FT - want[:, None, None]).max()
        o.close()
        errs.append((eFR, eFT))
    errs = np.array(errs)
    assert np.all(errs[-1] < 5e-3), errs
    ratio = errs[:-1] / errs[1:]
    assert np.all(ratio > 3.0), ratio


def test_wall_d2z_converges_on_the_tanh_grid():
    """The one-sided D2z of BX on both walls (through ORM * D2z(BX) in WW1,
    every other term zero) converges to d2BX/dz2 of a smooth profile, at the
    first order of the one-sided second difference."""
    errs = []
    for n in (17, 33, 65):
        desc = _tanh_desc(n, lMag=True)
        c = desc.consts
        zeros = np.zeros(desc.shape)
        b = lambda z: 0.2 * np.cos(1.7 * z) + 0.1 * z ** 3
        d2b = lambda z: -0.2 * 1.7 ** 2 * np.cos(1.7 * z) + 0.6 * z
        st = {"RO": np.ones(desc.shape), "TT": np.ones(desc.shape), "RU": zeros, "RV": zeros,
              "RW": zeros, "BX": _zfield(desc, b), "BY": zeros, "BZ": zeros}
        o = om.Oracle(desc)
        o.set_state(st)
        o.fluxes()
        W1 = o.get("WW1")
        z = desc.metrics.ZEE
        e = max(abs(W1[k, 1:-1, 1:-1] - c["ORM"] * d2b(z[k])).max() for k in (1, n))
        o.close()
        errs.append(e)
    errs = np.array(errs)
    ratio = errs[:-1] / errs[1:]
    assert np.all(ratio > 1.7), (errs, ratio)


def _cfl_state(desc, rng, bscale=1.0):
    nz, ny, nx = desc.shape
    return {"RO": 1.0 + rng.uniform(size=(nz, ny, nx)),
            "TT": 1.0 + rng.uniform(size=(nz, ny, nx)),
            "RU": rng.uniform(-1, 1, (nz, ny, nx)), "RV": rng.uniform(-1, 1, (nz, ny, nx)),
            "RW": rng.uniform(-1, 1, (nz, ny, nx)), "BX": bscale * rng.uniform(-1, 1, (nz, ny, nx)),
            "BY": bscale * rng.uniform(-1, 1, (nz, ny, nx)), "BZ": bscale * rng.uniform(-1, 1, (nz, ny, nx))}


def _dd(desc):
    m = desc.metrics
    return np.minimum(np.minimum(m.DDY[None, 1:-1, None], m.DDZ[1:-1, None, None]),
                      m.DDX[None, None, 1:-1])


@pytest.mark.parametrize("RM", [25.0, 1e-3])
def test_cfl_limits_on_the_tanh_grid(RM):
    """TSTEP (3dmhdsub.h:70-105) on the stretched grid: the thermal limit
    0.5 DD^2 REPR CV RO / RKAPA, the viscous limit 0.375 DD^2 RE RO and the
    magnetic limit 0.5 DD^2 RM each equal an independent minimum; with RM
    small the magnetic limit binds and DT = SFF * 0.5 * DDmin^2 * RM."""
    rng = np.random.default_rng(8)
    par = _par(RM)
    nz = 11 + 2
    rk = 1.0 + 0.5 * np.linspace(0, 1, nz)
    cfg = RunConfig(nPx=12, nPy=10, nPz=11, lMag=True, izCon=1, iTCon=4, par=par,
                    zgrid=("tanh", BETA), kappa=(rk, np.zeros(nz)))
    desc = Descriptor(cfg)
    c = desc.consts
    st = _cfl_state(desc, rng)
    o = om.Oracle(desc)
    o.set_state(st)
    dt, w, _ = o.cfl()
    o.close()
    sl = (slice(1, -1), slice(1, -1), slice(1, -1))
    rho = st["RO"][sl]
    DD = _dd(desc)
    thermal = 0.5 * DD ** 2 * c["REPR"] * c["CV"] * rho / rk[1:-1, None, None]
    viscous = 0.375 * DD ** 2 * c["RE"] * rho
    magnetic = 0.5 * DD.min() ** 2 * RM
    assert w[1] == pytest.approx(thermal.min(), rel=1e-14)
    assert w[2] == pytest.approx(viscous.min(), rel=1e-14)
    assert w[3] == pytest.approx(magnetic, rel=1e-14)
    assert dt == c["SFF"] * min(w)
    if RM < 1:
        assert w[3] < min(w[0], w[1], w[2])
        assert dt == pytest.approx(c["SFF"] * 0.5 * DD.min() ** 2 * RM, rel=1e-14)
