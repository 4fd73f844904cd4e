"""The oracle against what the mathematics fixes (CPU).

The published known answers pin the hydro and lRot paths of the oracle; the
magnetic block, the wall one-sided stencils and the BCON variants have no
published data, so they are checked here against independent statements of
the same physics:

* polynomial fields on a uniform grid, where second-order central differences
  are exact (D1 for quadratics, D2 for cubics): every flux equals its
  continuum expression -- continuity, momentum with the full stress tensor,
  energy with the viscous dissipation written as a sum of squared
  differences, induction as (B.grad)u - B div u - (u.grad)B + eta lap B,
  Lorentz force as (curl B) x B, Joule heating as |curl B|^2;
* telescoping sums: on a periodic uniform grid the bulk sums of the
  conservative fluxes reduce to the z-boundary terms;
* BCON: stress-free extrapolation, fixed-flux one-sided derivative, fixed
  temperatures;
* CFL: the advective limit equals DD / (|u| + c_fast).
"""
import numpy as np
import pytest

from mhd3d.desc import Descriptor, RunConfig, default_par
from oracle import oracle as om

NPX, NPY, NPZ = 12, 10, 9


def _cfg(lMag=True, lRot=False, izCon=1, iTCon=4, TP=0.0, TB=0.0):
    par = default_par()
    par[0] = 40.0          # RE
    par[1] = 0.7           # PR
    par[2] = 3.0           # THETA
    par[3] = 8.0           # GRAV
    par[4] = 2.0           # RY
    par[5] = 0.3           # ANG
    par[6] = 25.0          # RM
    par[7] = 1.5           # BETA
    par[11] = TP
    par[16], par[17], par[18] = 2.0, 1.5, 1.0
    return RunConfig(nPx=NPX, nPy=NPY, nPz=NPZ, lMag=lMag, lRot=lRot, izCon=izCon,
                     iTCon=iTCon, par=par, TB=TB)


def _coords(desc):
    """Physical coordinates implied by the uniform-grid stencil: D1x f =
    (f[i+1]-f[i-1]) * dxxdx * hx = (f(X+D)-f(X-D)) / (2D), D = 1/(2 hx dxxdx)."""
    c = desc.consts
    m = desc.metrics
    dx = 1.0 / (2.0 * c["hx"] * m.dxxdx[1])
    dy = 1.0 / (2.0 * c["hy"] * m.dyydy[1])
    dz = 1.0 / (2.0 * c["hz"] * m.dzzdz[1])
    nz, ny, nx = desc.shape
    Z, Y, X = np.meshgrid(np.arange(nz) * dz, np.arange(ny) * dy, np.arange(nx) * dx,
                          indexing="ij")
    return X, Y, Z


class Poly:
    """f = a + g.x + x^T H x (H symmetric) with analytic derivatives."""

    def __init__(self, rng, lin_scale=1.0, quad_scale=0.0, const=0.0):
        self.a = const + rng.uniform(-1, 1)
        self.g = rng.uniform(-1, 1, 3) * lin_scale
        H = rng.uniform(-1, 1, (3, 3)) * quad_scale
        self.H = 0.5 * (H + H.T)

    def __call__(self, X, Y, Z):
        P = np.stack([X, Y, Z])
        lin = self.g[0] * X + self.g[1] * Y + self.g[2] * Z
        quad = np.einsum("i...,ij,j...->...", P, self.H, P)
        return self.a + lin + quad

    def grad(self, X, Y, Z):
        P = np.stack([X, Y, Z])
        return self.g[:, None, None, None] + 2.0 * np.einsum("ij,j...->i...", self.H, P)

    def lap(self):
        return 2.0 * np.trace(self.H)


def _poly_case(seed, w_zero_plane=None):
    rng = np.random.default_rng(seed)
    desc = Descriptor(_cfg())
    X, Y, Z = _coords(desc)
    rho0 = 1.3
    T = Poly(rng, 0.3, 0.2, const=2.0)          # quadratic temperature
    u = [Poly(rng, 0.05) for _ in range(3)]     # linear velocity
    if w_zero_plane is not None:
        # impenetrable wall: w = alpha (z - z_wall) vanishes on that wall
        zw = Z[w_zero_plane, 0, 0]
        u[2].g = np.array([0.0, 0.0, u[2].g[2]])
        u[2].a = -u[2].g[2] * zw
    B = [Poly(rng, 0.02, 0.01) for _ in range(3)]  # quadratic field
    st = {
        "RO": np.full(X.shape, rho0),
        "TT": T(X, Y, Z),
        "RU": rho0 * u[0](X, Y, Z), "RV": rho0 * u[1](X, Y, Z), "RW": rho0 * u[2](X, Y, Z),
        "BX": B[0](X, Y, Z), "BY": B[1](X, Y, Z), "BZ": B[2](X, Y, Z),
    }
    o = om.Oracle(desc)
    o.set_state(st)                              # no periodic refill: polynomial ghosts
    o.fluxes()
    F = {n: o.get(n) for n in ("FU", "FV", "FW", "FR", "FT", "WW1", "WW2", "WW3")}
    return desc, (X, Y, Z), rho0, T, u, B, st, F


def _expected(desc, XYZ, rho0, T, u, B):
    c = desc.consts
    X, Y, Z = XYZ
    G = np.array([ui.g for ui in u])             # G[i, j] = d u_i / d x_j (constant)
    U = np.stack([ui(X, Y, Z) for ui in u])
    Bv = np.stack([b(X, Y, Z) for b in B])
    dB = np.stack([b.grad(X, Y, Z) for b in B])  # dB[i, j] = d B_i / d x_j
    Tv = T(X, Y, Z)
    dT = T.grad(X, Y, Z)
    div = np.trace(G)
    ORE, OCV, GAMMA = c["ORE"], c["OCV"], c["GAMMA"]
    C43 = c["C43"]
    # continuity and momentum (linear u: no viscous term survives; rho const)
    FR = -rho0 * div * np.ones_like(X)
    adv = np.stack([rho0 * sum(G[i, j] * U[j] + U[i] * G[j, j] for j in range(3)) for i in range(3)])
    gradP = rho0 * dT
    Fm = -gradP - adv
    Fm[2] += c["GRAV"] * rho0
    # Lorentz force OBETA (curl B) x B
    J = np.stack([dB[2, 1] - dB[1, 2], dB[0, 2] - dB[2, 0], dB[1, 0] - dB[0, 1]])
    Fm += c["OBETA"] * np.cross(J, Bv, axis=0)
    # energy: -u.grad T + kappa lap T / rho + heating - T div u + Joule
    kap = c["OCV_REPR"] if "OCV_REPR" in c else OCV / c["REPR"]
    FT = -(U * dT).sum(0) + kap * T.lap() / rho0
    e = G
    diag = (e[0, 0] - e[1, 1]) ** 2 + (e[1, 1] - e[2, 2]) ** 2 + (e[2, 2] - e[0, 0]) ** 2
    phi = C43 * 0.5 * diag + (e[0, 1] + e[1, 0]) ** 2 + (e[0, 2] + e[2, 0]) ** 2 + (e[1, 2] + e[2, 1]) ** 2
    FT = FT + OCV * ORE * phi / rho0 - OCV * Tv * div
    FT = FT + c["OBETA"] * c["ORM"] * OCV * (J ** 2).sum(0) / rho0
    # induction (B.grad)u - B div u - (u.grad)B + ORM lap B
    Fb = np.stack([
        sum(Bv[j] * G[i, j] for j in range(3)) - Bv[i] * div
        - sum(U[j] * dB[i, j] for j in range(3)) + c["ORM"] * B[i].lap()
        for i in range(3)])
    return {"FU": Fm[0], "FV": Fm[1], "FW": Fm[2], "FR": FR, "FT": FT,
            "WW1": Fb[0], "WW2": Fb[1], "WW3": Fb[2]}


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mhd_fluxes_equal_continuum_for_polynomial_fields(seed):
    desc, XYZ, rho0, T, u, B, st, F = _poly_case(seed)
    E = _expected(desc, XYZ, rho0, T, u, B)
    # interior points two or more cells from the periodic wrap; all planes
    # between the walls (walls are checked below for the terms they keep)
    sl = (slice(2, -2), slice(2, -2), slice(2, -2))
    for name in ("FR", "FU", "FV", "FW", "FT", "WW1", "WW2", "WW3"):
        g, e = F[name][sl], E[name][sl]
        scale = max(np.abs(e).max(), 1e-30)
        err = np.abs(g - e).max() / scale
        assert err < 1e-9, (name, err)


def test_wall_one_sided_stencils():
    """At the walls the reference replaces D1z(RW) in the continuity equation
    and D1z(w) in the induction by (4 f1 - f2)-type one-sided formulas that
    assume f = 0 on the wall (RW = 0 after BCON, 3dmhdsub.h:211-226,572-584),
    and D2z(BX), D2z(BY) by one-sided formulas exact for cubics (:640-687).
    With w vanishing on the wall and polynomial ghosts, every wall flux must
    equal its continuum value."""
    for zw in (1, NPZ):
        desc, XYZ, rho0, T, u, B, st, F = _poly_case(7, w_zero_plane=zw)
        E = _expected(desc, XYZ, rho0, T, u, B)
        sl = (zw, slice(2, -2), slice(2, -2))
        for name in ("FR", "WW1", "WW2", "WW3", "FU", "FV", "FT"):
            g, e = F[name][sl], E[name][sl]
            err = np.abs(g - e).max() / max(np.abs(e).max(), 1e-30)
            assert err < 1e-9, (zw, name, err)


def test_d2z_one_sided_is_exact_for_cubics():
    """The one-sided D2z of 3dmhdsub.h:640-654 at the walls reproduces the
    second derivative of a cubic exactly (checked through ORM*D2z(BX) in WW1
    with every other term zero)."""
    desc = Descriptor(_cfg())
    X, Y, Z = _coords(desc)
    c = desc.consts
    zs = Z - Z[NPZ // 2]
    st = {"RO": np.ones(X.shape), "TT": np.ones(X.shape),
          "RU": np.zeros(X.shape), "RV": np.zeros(X.shape), "RW": np.zeros(X.shape),
          "BX": 0.3 * zs ** 3 - 0.2 * zs ** 2 + 0.1 * zs + 0.05,
          "BY": np.zeros(X.shape), "BZ": np.zeros(X.shape)}
    o = om.Oracle(desc)
    o.set_state(st)
    o.fluxes()
    W1 = o.get("WW1")
    for zw in (1, NPZ):
        expect = c["ORM"] * (6 * 0.3 * zs[zw, 2:-2, 2:-2] - 2 * 0.2)
        got = W1[zw, 2:-2, 2:-2]
        assert np.abs(got - expect).max() < 1e-9 * max(1.0, np.abs(expect).max())


def test_conservative_fluxes_telescope():
    """On a periodic uniform grid the x and y parts of FR and FU telescope:
    their bulk sums over planes 2..nPz-1 reduce to z-boundary terms."""
    rng = np.random.default_rng(5)
    desc = Descriptor(_cfg(lMag=False))
    nz, ny, nx = desc.shape
    c = desc.consts
    st = {"RO": 1.0 + 0.1 * rng.uniform(size=(nz, ny, nx)),
          "TT": 1.0 + 0.1 * rng.uniform(size=(nz, ny, nx)),
          "RU": 0.1 * rng.uniform(-1, 1, (nz, ny, nx)),
          "RV": 0.1 * rng.uniform(-1, 1, (nz, ny, nx)),
          "RW": 0.1 * rng.uniform(-1, 1, (nz, ny, nx))}
    for a in st.values():
        a[:, 1:-1, 0] = a[:, 1:-1, -2]
        a[:, 1:-1, -1] = a[:, 1:-1, 1]
        a[:, 0, :] = a[:, -2, :]
        a[:, -1, :] = a[:, 1, :]
    o = om.Oracle(desc)
    o.set_state(st)
    o.fluxes()
    FR = o.get("FR")
    RW = st["RW"]
    dz = desc.metrics.dzzdz[2]
    bulk = FR[2:NPZ, 1:-1, 1:-1].sum()
    edge = -(RW[NPZ, 1:-1, 1:-1] + RW[NPZ - 1, 1:-1, 1:-1] - RW[1, 1:-1, 1:-1]
             - RW[2, 1:-1, 1:-1]).sum() * dz * c["hz"]
    assert abs(bulk - edge) < 1e-10 * np.abs(FR).sum()


@pytest.mark.parametrize("izCon,iTCon", [(0, 0), (1, 1), (1, 3), (0, 4)])
def test_bcon_variants(izCon, iTCon):
    rng = np.random.default_rng(11)
    desc = Descriptor(_cfg(lMag=True, izCon=izCon, iTCon=iTCon, TP=0.1 if iTCon == 4 else 0.0,
                           TB=7.5))
    X, Y, Z = _coords(desc)
    c = desc.consts
    zz = Z
    # stress-free profiles: quadratic in z with zero slope at each wall
    u_top = 0.2 + 0.3 * (zz - zz[1]) ** 2
    T = 2.0 + 0.4 * zz + 0.3 * zz ** 2
    st = {"RO": 1.0 + 0.2 * zz, "TT": T.copy(), "RU": (1.0 + 0.2 * zz) * u_top,
          "RV": (1.0 + 0.2 * zz) * 2 * u_top, "RW": rng.uniform(size=X.shape),
          "BX": rng.uniform(size=X.shape), "BY": rng.uniform(size=X.shape),
          "BZ": rng.uniform(size=X.shape)}
    o = om.Oracle(desc)
    o.set_state(st)
    o.bcon()
    RU, RW, TT, BZ, RO = o.get("RU"), o.get("RW"), o.get("TT"), o.get("BZ"), o.get("RO")
    top, bot = 1, NPZ
    inner = (slice(1, -1), slice(1, -1))
    # stress-free top: u_wall = (4 u1 - u2)/3 (zero one-sided slope), times rho
    u1, u2 = u_top[top + 1][inner], u_top[top + 2][inner]
    assert np.allclose(RU[top][inner] / RO[top][inner], (4 * u1 - u2) / 3, rtol=1e-6)
    assert np.all(RW[top][inner] == 0.0) and np.all(RW[bot][inner] == 0.0)
    assert np.all(BZ[top][inner] == 0.0) and np.all(BZ[bot][inner] == 0.0)
    dzc = desc.metrics.dzzdz
    if iTCon == 4:
        xs = np.arange(NPX + 2)[None, :]
        want = np.where(xs < (NPX + 2) // 2, 1.0 - c["TP"], 1.0)
        assert np.array_equal(TT[top][1:-1, 1:-1], np.broadcast_to(want, TT[top].shape)[1:-1, 1:-1])
    elif iTCon == 0:
        assert np.all(TT[top][inner] == 1.0)
    elif iTCon == 3:
        assert np.array_equal(TT[top][inner], T[top][inner])
    elif iTCon == 1:
        # second-order one-sided derivative at the top equals THETA (code units)
        t0, t1, t2 = TT[top][inner], TT[top + 1][inner], TT[top + 2][inner]
        dz_code = c["ORZF"] / dzc[top]
        slope = (-3 * t0 + 4 * t1 - t2) / (2 * dz_code)
        assert np.allclose(slope, c["THETA"], rtol=1e-6)
    if izCon == 0:
        assert np.all(TT[bot][inner] == 7.5)
    else:
        t0, t1, t2 = TT[bot][inner], TT[bot - 1][inner], TT[bot - 2][inner]
        dz_code = c["ORZF"] / dzc[bot]
        slope = (3 * t0 - 4 * t1 + t2) / (2 * dz_code)
        assert np.allclose(slope, -c["THETA"] / desc.metrics.RKAPA[bot], rtol=1e-6)


@pytest.mark.parametrize("mag", [False, True])
def test_cfl_advective_limit(mag):
    rng = np.random.default_rng(3)
    desc = Descriptor(_cfg(lMag=mag))
    nz, ny, nx = desc.shape
    c = desc.consts
    st = {"RO": 1.0 + rng.uniform(size=(nz, ny, nx)),
          "TT": 1.0 + rng.uniform(size=(nz, ny, nx)),
          "RU": rng.uniform(-1, 1, (nz, ny, nx)), "RV": rng.uniform(-1, 1, (nz, ny, nx)),
          "RW": rng.uniform(-1, 1, (nz, ny, nx)), "BX": rng.uniform(-1, 1, (nz, ny, nx)),
          "BY": rng.uniform(-1, 1, (nz, ny, nx)), "BZ": rng.uniform(-1, 1, (nz, ny, nx))}
    o = om.Oracle(desc)
    o.set_state(st)
    dt, w, _ = o.cfl()
    sl = (slice(1, -1), slice(1, -1), slice(1, -1))
    rho = st["RO"][sl]
    speed = np.sqrt(st["RU"][sl] ** 2 + st["RV"][sl] ** 2 + st["RW"][sl] ** 2) / rho
    cs2 = c["GAMMA"] * st["TT"][sl]
    if mag:
        cs2 = cs2 + c["OBETA"] * (st["BX"][sl] ** 2 + st["BY"][sl] ** 2 + st["BZ"][sl] ** 2) / rho
    m = desc.metrics
    DD = np.minimum(np.minimum(m.DDY[None, 1:-1, None], m.DDZ[1:-1, None, None]), m.DDX[None, None, 1:-1])
    adv = DD / (speed + np.sqrt(cs2))
    assert abs(w[0] - adv.min()) <= 1e-13 * adv.min()
    lims = [w[0], w[1], w[2]] + ([w[3]] if mag else [])
    assert dt == c["SFF"] * min(lims)
