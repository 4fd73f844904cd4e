"""The product's Static() (mhd3d_static_state, host code) against the oracle's
restatement of the same routine, which the reference's published iteration-9
answers pin (tests/test_oracle_known_answers.py): bit-identical interiors,
including the default 4x6-rank run and a run whose planes do not divide
evenly among the z ranks."""
import os

import numpy as np
import pytest

from known_answers import default_config, default_state
from mhd3d import _lib as L
from mhd3d.initial import reference_static_state
from oracle import oracle as om

pytestmark = pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="libmhd3d.so not built")

INNER = (slice(None), slice(1, -1), slice(1, -1))


def _check(cfg, NP, NPE_Y, NPE_Z):
    ref = default_state(cfg, om, NP=NP, NPE_Y=NPE_Y, NPE_Z=NPE_Z)
    got = reference_static_state(cfg, NP, NPE_Y, NPE_Z)
    for k in ref:
        assert np.array_equal(got[k][INNER], ref[k][INNER]), k
    # x / y ghosts are periodic copies
    TT = got["TT"]
    assert np.array_equal(TT[:, 1:-1, 0], TT[:, 1:-1, -2]) and np.array_equal(TT[:, 0, :], TT[:, -2, :])


def test_default_rank_layout():
    cfg = default_config()          # 256 x 256 x 252 of the 4 x 6 rank run (256 global planes)
    _check(cfg, 256, 4, 6)


def test_small_uneven_layout():
    from dataclasses import replace
    cfg = replace(default_config(), nPx=16, nPy=12, nPz=18)   # 20 planes on 3 z ranks of 6
    _check(cfg, 20, 2, 3)


def _check_flags(cfg, NP, NPE_Y, NPE_Z):
    ref, rprof = default_state(cfg, om, NP=NP, NPE_Y=NPE_Y, NPE_Z=NPE_Z, with_profiles=True)
    got, gprof = reference_static_state(cfg, NP, NPE_Y, NPE_Z, with_profiles=True)
    for k in ref:
        assert np.array_equal(got[k][INNER], ref[k][INNER]), k
    if cfg.lRem:
        for k in rprof:
            assert np.array_equal(np.asarray(gprof[k]), np.asarray(rprof[k])), k
    return got, gprof


def _small(**kw):
    from dataclasses import replace
    return replace(default_config(), nPx=16, nPy=12, nPz=18, **kw)


@pytest.mark.parametrize("ampb", [0.0, 0.02])
def test_magnetic_layer(ampb):
    cfg = _small(lMag=True)
    par = cfg.par.copy()
    par[7] = 1.5                     # BETA
    par[20], par[21], par[22] = ampb, 0.3, 0.6   # AMPB, BFH, BZP
    cfg.par = par
    got, _ = _check_flags(cfg, 20, 2, 3)
    if ampb == 0.0:
        assert not got["BX"].any() and not got["BY"].any() and not got["BZ"].any()
        return
    # Gaussian of half width BFH/2 about BZP: BX = AMPB * 2^(-4 (z-BZP)^2 / BFH^2) (float ln 2);
    # each rank's first plane (1, 7, 13) keeps B = 0 (3dmhdphsic.h:684)
    nrz, zmax = 6, par[18]
    for z in range(1, 19):           # the planes of the 18-plane domain
        bx = got["BX"][z, 3, 4]
        if (z - 1) % nrz == 0:
            assert bx == 0.0
            continue
        k = z - 1 if (z - 1) // nrz < 2 else z - 1 + 2
        zee = k / 19 * zmax
        assert abs(bx - 0.02 * 2.0 ** (-4 * (zee - 0.6) ** 2 / 0.3 ** 2)) <= 1e-9
    # magnetic pressure balance of the layer, before the RAN2 perturbation
    # (AMPT = 0): RO*TT + BX^2*OBETA = the plain state's RO*TT
    quiet_cfg, plain_cfg = _small(lMag=True), _small()
    quiet_cfg.par = par.copy()
    plain_cfg.par = plain_cfg.par.copy()
    quiet_cfg.par[19] = plain_cfg.par[19] = 0.0
    quiet = reference_static_state(quiet_cfg, 20, 2, 3)
    plain = reference_static_state(plain_cfg, 20, 2, 3)
    ptot = quiet["RO"] * quiet["TT"] + quiet["BX"] ** 2 * (2.0 / 1.5)
    base = plain["RO"] * plain["TT"]
    assert np.allclose(ptot[1:-1, 1:-1, 1:-1], base[1:-1, 1:-1, 1:-1], rtol=1e-12, atol=0)


def test_shear_start():
    cfg = _small(lShr=True)
    got, _ = _check_flags(cfg, 20, 2, 3)
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    assert np.all(got["TT"][inner] == 1.0) and np.all(got["RO"][inner] == 1.0)
    amp = cfg.par[19]
    rv = got["RV"][inner]
    assert rv.std() > 0.2 * amp and np.all(np.abs(rv) <= 0.5 * amp + 1e-15)
    # the same RAN2 draws that multiply TT without lShr
    plain = reference_static_state(_small(), 20, 2, 3)
    ratio = plain["TT"][inner] / plain["TT"][inner].mean(axis=(1, 2), keepdims=True)
    assert np.corrcoef(rv.ravel(), ratio.ravel())[0, 1] > 0.99


def _rem_cfg():
    cfg = _small(lRem=True)
    cfg.par = cfg.par.copy()
    cfg.par[18] = 1.0            # zMax < 2: see test_rempel_background_needs_zmax_below_two
    return cfg


def test_rempel_background_profiles():
    got, prof = _check_flags(_rem_cfg(), 20, 2, 3)
    assert got["TT"].shape[0] == len(prof["RKAPA"])
    assert prof["TU"] == got["TT"][0, 5, 5]


def test_rempel_background_needs_zmax_below_two():
    # the integration starts at x = 2; a plane at height 2 (zMax = 2, the
    # default) gives a zero step, where the reference's RKQC exits
    # ("Stepsize not significant in RKQC.", 3dmhdphsic.h:485-488)
    with pytest.raises(L.MhdError) as ei:
        reference_static_state(_small(lRem=True), 20, 2, 3)
    assert ei.value.code == L.EINVAL
    with pytest.raises(RuntimeError):
        om.rempel_profile(20, 5.0, 15.0, 2.0)


def test_rempel_background_solves_its_ode():
    """The oracle's ODEINT of derivs1 against an independent high-accuracy
    solve (scipy, rtol 1e-12) of dT/dx = min-type gradient, dP/dx = P/T from
    x = 2, T = P = 1; EPS = 1e-8 per step bounds the difference."""
    from scipy.integrate import solve_ivp

    def rhs(x, y):
        T, P = y
        rho = P / T
        krad = rho + (1.25 - rho) * np.exp(-5.0 * x)
        nrad = 0.4 / krad
        return [nrad if nrad <= 0.41 else 0.41, P / T]

    NPz, zmax, theta, repr_ = 20, 1.0, 5.0, 75.0 * 0.2
    T, R, RK, DRK, walls = om.rempel_profile(NPz, theta, repr_, zmax)
    for k in range(NPz):
        x = k / (NPz - 1) * zmax
        sol = solve_ivp(rhs, (2.0, x), [1.0, 1.0], rtol=1e-12, atol=1e-14, method="DOP853")
        Te, Pe = sol.y[:, -1]
        assert abs(T[k] - Te) <= 1e-6 * Te
        assert abs(R[k] - Pe / Te) <= 1e-6 * Pe / Te
        kap = R[k] + (1.25 - R[k]) * np.exp(-5.0 * x)
        assert abs(RK[k] - kap * 8.07 * theta * repr_) <= 1e-6 * RK[k]
    # DRK is dRK/dz on the uniform grid (one-sided at the ends): second order
    z = np.arange(NPz) / (NPz - 1) * zmax
    dk = np.gradient(RK, z, edge_order=2)
    assert np.abs(DRK - dk).max() <= 2e-2 * np.abs(dk).max()
    assert walls[0] == T[0]


def test_lshr_with_lrem_refused():
    cfg = _small(lShr=True, lRem=True)
    with pytest.raises(L.MhdError) as ei:
        reference_static_state(cfg, 20, 2, 3)
    assert ei.value.code == L.EUNSUPPORTED
