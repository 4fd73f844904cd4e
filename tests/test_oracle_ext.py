"""The oracle's off-default branches against what the mathematics fixes (CPU).

No published data covers iD, lRem, lShr, iTCon 2 or the TP plume, so each is
pinned by an independent statement of the same operator:

* density diffusion (iD 1/2): a density r0(z) + s(z) cos(kx) has plane mean
  r0(z); the diffusion term is then the discrete x-Laplacian eigenvalue
  -4 sin^2(pi/nPx) times s cos(kx) and, in z, the exact second difference of
  the quadratic s(z) -- so a missing mean removal, a wrong weight 1/ROM or a
  wrong coefficient each fail; the iD 2 weight is a sum of two Gaussians whose
  half-maximum is written here as a power of two;
* lRem conduction: for T = t0(z) + a(z) cos(kx) at rest the temperature flux is
  OCV/REPR/rho (Lap T' + RKAPA t0'' + DKAPA t0') with the mean t0 and the
  fluctuation T' treated apart; the CFL conduction limit takes 1 + RKAPA;
  the lRem walls keep a linear temperature when DZTU/DZTB are the
  profile's own one-sided differences (what Static() stores);
* lShr: FT is carried over between calls (it doubles on a second call at
  rest) and holds only conduction with motion; the walls move at +-rho;
* iTCon 2: the heat loss is -amp/rho times Gaussians that halve at
  (x-XP)^2 = HH^3/4; at TIMT - TC = QFH its amplitude is 0.8 TP OCV
  (tanh(ln 2) = 3/5);
* TP plume: at TIMT = TC ln 2, TPR = TP + (1-TP)/2 (exp(-ln 2) = 1/2); the
  centre takes the Gaussian's peak and far columns the plain wall value.

Tolerances: 1e-10 relative for exact identities; 1e-7 where the reference
rounds ln 2 to float (the Gaussian exponents).
"""
import math

import numpy as np
import pytest

from mhd3d.desc import Descriptor, RunConfig, default_par
from oracle import oracle as om
from test_oracle_math import _coords

NPX, NPY, NPZ = 16, 12, 10


def _par(TP=0.0, TC=0.0, DH=0.0, DV=0.0, HH=0.5, QFH=0.25, XP=0.0, YP=0.0, ZP=0.0):
    par = default_par()
    par[0] = 40.0          # RE
    par[1] = 0.7           # PR
    par[2] = 3.0           # THETA
    par[3] = 8.0           # GRAV
    par[11], par[12] = TP, TC
    par[13], par[14], par[15] = XP, YP, ZP
    par[16], par[17], par[18] = 2.0, 1.5, 1.0
    par[52], par[54] = QFH, HH
    par[55], par[56] = DH, DV
    return par


def _desc(**kw):
    par = _par(**{k: kw.pop(k) for k in list(kw) if k in ("TP", "TC", "DH", "DV", "HH", "QFH", "XP", "YP", "ZP")})
    base = dict(nPx=NPX, nPy=NPY, nPz=NPZ, izCon=1, iTCon=4, par=par)
    base.update(kw)
    return Descriptor(RunConfig(**base))


def _cosx(desc):
    nz, ny, nx = desc.shape
    th = 2 * np.pi * (np.arange(nx) - 1) / NPX
    return np.broadcast_to(np.cos(th), (nz, ny, nx)).copy()


def _lam_x(desc):
    """D2x of cos(kx) over cos(kx) on the uniform periodic grid."""
    m, c = desc.metrics, desc.consts
    return -4.0 * math.sin(math.pi / NPX) ** 2 * m.dxxdx[1] ** 2 * c["h2x"]


def _rest_state(desc, RO, TT):
    z = np.zeros(desc.shape)
    return {"RU": z, "RV": z.copy(), "RW": z.copy(), "RO": RO, "TT": TT}


def _fluxes(desc, st, names=("FR", "FT"), calls=1):
    o = om.Oracle(desc)
    o.set_state(st)
    o.communicate()
    for _ in range(calls):
        o.fluxes()
    return {n: o.get(n) for n in names}, o


def _close(a, b, tol, region=(slice(1, -1), slice(1, -1), slice(1, -1))):
    a, b = a[region], b[region]
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= tol * scale, (np.abs(a - b).max(), scale)


# ---------------------------------------------------------------- iD


def _diffusion_case(iD, DH, DV):
    desc = _desc(iD=iD, DH=DH, DV=DV)
    X, Y, Z = _coords(desc)
    r0 = 1.2 + 0.3 * Z - 0.2 * Z ** 2           # mean profile (curved: its D2z must not enter)
    s = 0.05 + 0.02 * Z + 0.03 * Z ** 2         # amplitude of the fluctuation
    RO = r0 + s * _cosx(desc)
    TT = np.full(desc.shape, 1.7)
    base = _desc(iD=0)
    F1, _ = _fluxes(desc, _rest_state(desc, RO, TT))
    F0, _ = _fluxes(base, _rest_state(base, RO, TT))
    return desc, (X, Y, Z), r0, s, F1["FR"] - F0["FR"]


@pytest.mark.parametrize("DH,DV", [(0.7, 0.0), (0.0, 0.9), (0.7, 0.9)])
def test_density_diffusion_iD1(DH, DV):
    desc, (X, Y, Z), r0, s, dFR = _diffusion_case(1, DH, DV)
    c, m = desc.consts, desc.metrics
    cosx = _cosx(desc)
    # D2z of s(z) = 0.05 + 0.02 Z + 0.03 Z^2 on the uniform grid: 0.06 (exact)
    lap_h = _lam_x(desc) * s * cosx                      # D2x + D2y (no y dependence)
    exp = (DH * c["ORE"]) * lap_h / r0 + (DV * c["ORE"]) * (0.06 * cosx) / r0
    _close(dFR, exp, 1e-10, (slice(2, -2), slice(1, -1), slice(1, -1)))
    # on the walls the ghost plane holds the plane two in: D2z(W) = 2 (W2 - W1) / dz^2
    dz = 1.0 / (2 * c["hz"] * m.dzzdz[1])
    for zw, zi in ((1, 2), (NPZ, NPZ - 1)):
        W1 = s[zw] * cosx[zw]
        W2 = s[zi] * cosx[zi]
        expw = (DH * c["ORE"]) * lap_h[zw] / r0[zw] + (DV * c["ORE"]) * (2 * (W2 - W1) / dz ** 2) / r0[zw]
        got = dFR[zw, 1:-1, 1:-1]
        want = expw[1:-1, 1:-1]
        assert np.abs(got - want).max() <= 1e-9 * max(np.abs(want).max(), 1e-12)


def test_density_diffusion_iD2_weights_are_gaussians():
    desc, (X, Y, Z), r0, s, dFR = _diffusion_case(2, 0.7, 0.0)
    c = desc.consts
    cosx = _cosx(desc)
    base = (0.7 * c["ORE"]) * _lam_x(desc) * s * cosx     # times wwz(z)
    zmax = c["zMax"]
    ZEE = desc.metrics.ZEE
    # exp(-4 ln2 z^2 / SGM^2), SGM = 1: half maximum at z = 1/2, i.e. 2^(-4 z^2)
    w = 2.0 ** (-4 * ZEE ** 2) + 2.0 ** (-4 * (ZEE - zmax) ** 2)
    exp = base * w[:, None, None]
    _close(dFR, exp, 1e-7)


# ---------------------------------------------------------------- lRem


def _rem_desc(iTCon=4, izCon=1, TU=0.0, DZTU=0.0, DZTB=0.0):
    nz = NPZ + 2
    zs = np.linspace(0, 1, nz)
    RK = 1.0 + 0.4 * zs ** 2                     # any conductivity profile
    DRK = 0.3 + 0.1 * zs
    return _desc(lRem=True, iTCon=iTCon, izCon=izCon, TU=TU, DZTU=DZTU, DZTB=DZTB, kappa=(RK, DRK))


def test_lrem_conduction_splits_mean_and_fluctuation():
    desc = _rem_desc()
    c, m = desc.consts, desc.metrics
    X, Y, Z = _coords(desc)
    rho = 1.3
    t0 = 2.0 + 0.5 * Z + 0.4 * Z ** 2            # mean profile: t0'' = 0.8, t0' = 0.5 + 0.8 Z
    a = 0.1 + 0.05 * Z                           # fluctuation amplitude (linear: no z curvature)
    TT = t0 + a * _cosx(desc)
    F, _ = _fluxes(desc, _rest_state(desc, np.full(desc.shape, rho), TT), names=("FT",))
    RK = m.RKAPA[:, None, None]
    DRK = m.DKAPA[:, None, None]
    lap_fluct = _lam_x(desc) * a * _cosx(desc)
    exp = (c["OCV"] / c["REPR"]) / rho * (lap_fluct + RK * 0.8 + DRK * (0.5 + 0.8 * Z))
    _close(F["FT"], exp, 1e-10)


def test_lrem_cfl_takes_one_plus_rkapa():
    desc = _rem_desc()
    c, m = desc.consts, desc.metrics
    o = om.Oracle(desc)
    rho = 1.1
    o.set_state(_rest_state(desc, np.full(desc.shape, rho), np.full(desc.shape, 1.5)))
    _, w, _ = o.cfl()
    DD = min(m.DDX[1:-1].min(), m.DDY[1:-1].min(), m.DDZ[1:-1].min())
    want = min(0.5 * DD * DD * c["REPR"] * c["CV"] * rho / (1.0 + m.RKAPA[z]) for z in range(1, NPZ + 1))
    assert abs(w[1] - want) <= 1e-12 * want


@pytest.mark.parametrize("iTCon", [1, 4])
def test_lrem_walls_keep_a_linear_profile(iTCon):
    # T = 2 + 0.7 Z; DZTU = T0 - 4/3 T1 + 1/3 T2, DZTB = T(n-1) - 4/3 T(n-2) + 1/3 T(n-3)
    # (3dmhdphsic.h:590-592) of that profile: the one-sided walls reproduce it
    probe = _rem_desc()
    X, Y, Z = _coords(probe)
    zc = Z[:, 0, 0]
    T = 2.0 + 0.7 * zc
    # Static() forms them with the float literals 4.0f/3.0f and 1.0f/3.0f, the
    # same roundings as C43 and C13, so the identity holds to the last bits
    c43, c13 = float(np.float32(4.0) / np.float32(3.0)), float(np.float32(1.0) / np.float32(3.0))
    DZTU = T[1] - c43 * T[2] + c13 * T[3]
    DZTB = T[NPZ] - c43 * T[NPZ - 1] + c13 * T[NPZ - 2]
    desc = _rem_desc(iTCon=iTCon, izCon=1, TU=T[1], DZTU=DZTU, DZTB=DZTB)
    o = om.Oracle(desc)
    TT = np.broadcast_to(T[:, None, None], desc.shape).copy()
    st = _rest_state(desc, np.ones(desc.shape), TT.copy())
    st["TT"][1] = -5.0
    st["TT"][NPZ] = -5.0
    o.set_state(st)
    o.bcon()
    got = o.get("TT")
    assert np.abs(got[1, 1:-1, 1:-1] - T[1]).max() < 1e-12
    assert np.abs(got[NPZ, 1:-1, 1:-1] - T[NPZ]).max() < 1e-12


# ---------------------------------------------------------------- lShr


def test_lshr_carries_ft_and_keeps_only_conduction():
    desc = _desc(lShr=True)
    plain = _desc()
    c = desc.consts
    X, Y, Z = _coords(desc)
    rho = 1.4
    TT = 2.0 + 0.3 * Z + 0.25 * Z ** 2
    st = _rest_state(desc, np.full(desc.shape, rho), TT)
    F1, _ = _fluxes(desc, st, names=("FT",), calls=1)
    F2, _ = _fluxes(desc, st, names=("FT",), calls=2)
    cond = (c["OCV"] / c["REPR"]) / rho * (0.5 * np.ones(desc.shape))   # RKAPA = 1, DKAPA = 0
    _close(F1["FT"], cond, 1e-10)
    _close(F2["FT"], 2 * cond, 1e-10)
    # with motion the lShr flux still holds conduction only; the plain one does not
    st_m = dict(st)
    st_m["RU"] = rho * (0.02 * X + 0.01 * Z)
    Fm, _ = _fluxes(desc, st_m, names=("FT",))
    Fp, _ = _fluxes(plain, st_m, names=("FT",))
    _close(Fm["FT"], cond, 1e-10)
    assert np.abs(Fp["FT"] - cond)[1:-1, 1:-1, 1:-1].max() > 1e-4


def test_lshr_with_lrem_conducts_the_pressure():
    """lShr with lRem (3dmhdsub.h:293-313): the TTM update and WW1 = TT - TTM
    sit inside `#if !lShr`, so lRem's conduction differentiates what WW1
    still holds, P = RO*TT (:267), and HRAD comes from the static TTM, zero.
    At rest with RO = rho and TT quadratic in z, P'' = rho*TT'' and FT =
    OCV/REPR * P''/rho = OCV/REPR * TT'' -- rho times the lShr-only value,
    with no RKAPA factor -- and it is carried over to the next call."""
    nz = NPZ + 2
    kap = (1.0 + 0.5 * np.linspace(0, 1, nz), 0.3 * np.ones(nz))
    desc = _desc(lShr=True, lRem=True, iTCon=1, izCon=1, TU=1.0, kappa=kap)
    c = desc.consts
    X, Y, Z = _coords(desc)
    rho = 1.4
    TT = 2.0 + 0.3 * Z + 0.25 * Z ** 2
    st = _rest_state(desc, np.full(desc.shape, rho), TT)
    F1, _ = _fluxes(desc, st, names=("FT",), calls=1)
    F2, _ = _fluxes(desc, st, names=("FT",), calls=2)
    cond = (c["OCV"] / c["REPR"]) * (0.5 * np.ones(desc.shape))
    _close(F1["FT"], cond, 1e-10)
    _close(F2["FT"], 2 * cond, 1e-10)


def test_lshr_walls_move():
    desc = _desc(lShr=True)
    o = om.Oracle(desc)
    rng = np.random.default_rng(3)
    st = {k: rng.uniform(0.5, 1.5, desc.shape) for k in ("RU", "RV", "RW", "RO", "TT")}
    o.set_state(st)
    o.bcon()
    for zw, sgn in ((1, 1.0), (NPZ, -1.0)):
        assert np.all(o.get("RU")[zw, 1:-1, 1:-1] == 0.0)
        assert np.all(o.get("RV")[zw, 1:-1, 1:-1] == sgn * st["RO"][zw, 1:-1, 1:-1])
        assert np.all(o.get("RW")[zw, 1:-1, 1:-1] == 0.0)


# ---------------------------------------------------------------- iTCon 2


def test_heat_loss_gaussian_and_amplitude():
    HH, QFH, TC, TP = 0.5, 0.25, 0.3, 0.2
    probe = _desc()
    X, Y, Z = _coords(probe)
    m = probe.metrics
    ix, iy, iz = 5, 4, 6                          # the source sits on this grid point
    XP, YP, ZP = m.EXX[ix], m.WYY[iy], m.ZEE[iz]
    desc = _desc(iTCon=2, TP=TP, TC=TC, HH=HH, QFH=QFH, XP=XP, YP=YP, ZP=ZP, TIMI=TC + QFH)
    c = desc.consts
    rho = 1.25
    F, _ = _fluxes(desc, _rest_state(desc, np.full(desc.shape, rho), np.full(desc.shape, 1.5)),
                   names=("FT",))
    amp = TP * c["OCV"] * 0.8 / rho              # 0.5 (1 + tanh(ln 2)) = 0.8
    assert abs(F["FT"][iz, iy, ix] + amp) <= 1e-7 * amp
    # the Gaussians exp(CLN d^2 / HH), CLN = -4 ln2 / HH^2: 2^(-4 d^2 / HH^3)
    gx = 2.0 ** (-4 * (m.EXX - XP) ** 2 / HH ** 3)
    gy = 2.0 ** (-4 * (m.WYY - YP) ** 2 / HH ** 3)
    gz = 2.0 ** (-4 * (m.ZEE - ZP) ** 2 / HH ** 3)
    exp = -amp * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    _close(F["FT"], exp, 1e-7)


def test_heat_loss_without_time_factor():
    probe = _desc()
    m = probe.metrics
    desc = _desc(iTCon=2, TP=0.2, TC=0.0, HH=0.5, XP=m.EXX[3], YP=m.WYY[3], ZP=m.ZEE[3])
    c = desc.consts
    F, _ = _fluxes(desc, _rest_state(desc, np.full(desc.shape, 2.0), np.full(desc.shape, 1.5)),
                   names=("FT",))
    assert abs(F["FT"][3, 3, 3] + 0.2 * c["OCV"] / 2.0) <= 1e-12


# ---------------------------------------------------------------- TP plume


@pytest.mark.parametrize("iTCon", [0, 1])
def test_plume_wall(iTCon):
    HH, TC, TP = 0.6, 0.4, 0.3
    probe = _desc()
    m = probe.metrics
    ix, iy = 6, 5
    XP, YP = m.EXX[ix], m.WYY[iy]
    desc = _desc(iTCon=iTCon, TP=TP, TC=TC, HH=HH, XP=XP, YP=YP, TIMI=TC * math.log(2.0))
    flat = _desc(iTCon=iTCon)
    c = desc.consts
    rng = np.random.default_rng(5)
    st = {k: rng.uniform(0.8, 1.2, desc.shape) for k in ("RU", "RV", "RW", "RO", "TT")}
    o = om.Oracle(desc)
    o.set_state(st)
    o.bcon()
    TT = o.get("TT")[1]
    TPR = TP + (1 - TP) * 0.5                   # exp(-ln 2) = 1/2
    g = (2.0 ** (-4 * (m.EXX - XP) ** 2 / HH ** 2))[None, :] * (2.0 ** (-4 * (m.WYY - YP) ** 2 / HH ** 2))[:, None]
    if iTCon == 0:
        exp = 1.0 - (1.0 - TPR) * g / HH ** 2
        assert np.abs(TT[1:-1, 1:-1] - exp[1:-1, 1:-1]).max() <= 1e-7
    else:
        # fixed-gradient wall whose flux THETA is lowered to TPR at the centre:
        # TT = C43 T1 - C13 T2 - C23 DZ (THETA - (THETA - TPR) g / HH^2)
        of = om.Oracle(flat)
        of.set_state(st)
        of.bcon()
        base = of.get("TT")[1]                   # THETA everywhere (DZ from float 1/(nPz-1))
        DZ = (1.0 / (NPZ - 1)) / m.dzzdz[1]
        exp = base + c["C23"] * DZ * (c["THETA"] - TPR) * g / HH ** 2
        assert np.abs(TT[1:-1, 1:-1] - exp[1:-1, 1:-1]).max() <= 1e-7
