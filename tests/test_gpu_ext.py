"""The off-default branches on the GPU against the oracle, bit for bit.

iD 1/2 (density diffusion, incl. the plane means it needs), lRem (the
conduction of TT - TTM, its walls and CFL; also together with iD 1, both
plane means in one pass), lShr (the carried-over FT, the
moving walls), iTCon 2 (heat loss with and without its time factor) and the
TP plume of iTCon 0/1 (3dmhdsub.h:229-326, 443-461, 720-873): 10 CFL steps,
single stages, fluxes() and BCON through the C ABI, compared over the
interior incl. both walls.  The plane means are summed in the reference's
order and the coordinate and time factors come from the same library
functions as the oracle's, so the bar is bit identity (and 1e-12 as the
stated tolerance of the FP64 form).
"""
import dataclasses

import numpy as np
import pytest

from gpu_common import bit_identical_fraction, relerr
from mhd3d import _lib as L
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.initial import stratified_config, stratified_state
from oracle import oracle as om

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _kappa(nz):
    zs = np.linspace(0.0, 1.0, nz)
    return 1.0 + 0.5 * zs ** 2, 1.0 * zs


def _cfg(n, case):
    lMag = case.endswith("_mhd")
    cfg = stratified_config(*n, lMag=lMag, lRot=lMag, Ra=3e4)
    par = cfg.par.copy()
    kw = {}
    if case.startswith("iDrem"):                 # density diffusion and lRem: both plane means
        kw = dict(iD=1, lRem=True, iTCon=1, izCon=1, TU=1.0, DZTU=-0.02, DZTB=0.03,
                  kappa=_kappa(n[2] + 2))
        par[55], par[56] = 0.5, 0.4
    elif case.startswith("iD1"):
        kw = dict(iD=1)
        par[55], par[56] = 0.5, 0.4
    elif case.startswith("iD2"):
        kw = dict(iD=2)
        par[55], par[56] = 0.3, 0.6
    elif case.startswith("iDh"):
        kw = dict(iD=1)
        par[55], par[56] = 0.5, 0.0
    elif case.startswith("lRem1"):
        kw = dict(lRem=True, iTCon=1, izCon=1, TU=1.0, DZTU=-0.02, DZTB=0.03,
                  kappa=_kappa(n[2] + 2))
    elif case.startswith("lRem0"):
        kw = dict(lRem=True, iTCon=0, izCon=0, TU=1.05, kappa=_kappa(n[2] + 2))
    elif case.startswith("lShrRem"):
        kw = dict(lShr=True, lRem=True, iTCon=1, izCon=1, TU=1.0, DZTU=-0.02, DZTB=0.03,
                  kappa=_kappa(n[2] + 2))
    elif case.startswith("lShr"):
        kw = dict(lShr=True)
    elif case.startswith("heat"):
        kw = dict(iTCon=2, TIMI=0.001)
        par[11], par[12] = 0.5, (0.002 if case.startswith("heatT") else 0.0)
        par[13], par[14], par[15] = 1.2, 0.9, 0.6
        par[52], par[54] = 0.01, 0.8
    elif case.startswith("plume"):
        it = int(case[5])
        kw = dict(iTCon=it, izCon=1 if it == 1 else 0)
        par[11], par[12] = 0.3, (0.004 if it == 0 else 0.0)
        par[13], par[14] = 1.4, 1.1
        par[54] = 0.9
    else:
        raise ValueError(case)
    return dataclasses.replace(cfg, par=par, **kw)


CASES = ["iD1", "iD2_mhd", "iDh_mhd", "lRem1", "lRem0", "lShr", "lShr_mhd", "lShrRem", "lShrRem_mhd",
         "heatT", "heat0_mhd", "plume0", "plume1", "iDrem", "iDrem_mhd"]


def _pair(n, case, seed=7):
    desc = Descriptor(_cfg(n, case))
    st = stratified_state(desc, seed=seed)
    dom = Domain(desc)
    dom.set_state(st)
    orc = om.Oracle(desc)
    orc.set_state(st)
    orc.communicate()
    return desc, dom, orc


def _same(dom, orc, what):
    for name, g in dom.state().items():
        c = orc.get(name)
        frac = bit_identical_fraction(g, c)
        assert frac == 1.0, (what, name, frac, relerr(g, c))
        assert relerr(g, c) <= TOL


@pytest.mark.parametrize("case", CASES)
def test_ten_steps_cfl(case):
    print(f"ext ten steps {case}", flush=True)
    desc, dom, orc = _pair((24, 20, 12), case)
    dts_c = [orc.step() for _ in range(10)]
    dt_g = dom.step(10)
    assert dt_g == dts_c[-1], (case, dt_g, dts_c[-1])
    _same(dom, orc, case)
    n, timt, timc = dom.clock()
    n_c, timt_c, timc_c = orc.clock()
    assert n == n_c == 10 and timc == timc_c and timt == timt_c


@pytest.mark.parametrize("case,n", [("iD1", (128, 16, 8)), ("lRem1", (128, 16, 8)),
                                    ("iD2_mhd", (136, 12, 8)), ("lRem0", (72, 10, 8)),
                                    ("iD1", (1, 600, 6)), ("lRem1", (3, 200, 6))])
def test_plane_means_wide_planes(case, n):
    """Planes of 2048 and 1632 values (whole groups of 512 summed by the
    pipelined path, chunks of 64 that span two rows), of 720 values (one
    whole group and a rest), and of 1-3 columns (a chunk of 64 spans 22-64
    rows; 600 values: one whole group and a rest): five CFL steps bit for
    bit."""
    print(f"ext wide planes {case} {n}", flush=True)
    desc, dom, orc = _pair(n, case, seed=11)
    dts_c = [orc.step() for _ in range(5)]
    assert dom.step(5) == dts_c[-1], case
    _same(dom, orc, (case, n))


@pytest.mark.parametrize("case", ["iD1", "iD2_mhd", "lRem1", "lShr", "heatT"])
def test_point_kernel_override(case, monkeypatch):
    """MHD3D_KERNEL=point (read when a context is created) runs the one-thread-
    per-point stage kernel instead of the z-streaming one: the same ten CFL
    steps bit for bit."""
    print(f"ext point kernel {case}", flush=True)
    monkeypatch.setenv("MHD3D_KERNEL", "point")
    desc, dom, orc = _pair((24, 20, 12), case, seed=13)
    dts_c = [orc.step() for _ in range(10)]
    assert dom.step(10) == dts_c[-1], case
    _same(dom, orc, case)


@pytest.mark.parametrize("case", ["iD2_mhd", "lRem1", "lShr_mhd", "lShrRem", "heatT", "plume0"])
def test_stages_and_fluxes(case):
    print(f"ext stages {case}", flush=True)
    desc, dom, orc = _pair((20, 16, 10), case, seed=3)
    fg = dom.fluxes()
    orc.fluxes()
    for name in L.FLUXES[: dom.nf]:
        c = orc.get(name)
        assert bit_identical_fraction(fg[name], c) == 1.0, (case, name, relerr(fg[name], c))
    if desc.cfg.lShr:
        # the oracle's FT now holds one call's worth; start both from a zero FT
        orc = om.Oracle(desc)
        orc.set_state(dom.state())
        orc.communicate()
        dom.close()
        dom = Domain(desc)
        dom.set_state(orc.state(mag=desc.cfg.lMag))
    dt, _, _ = orc.cfl()
    for k in (1, 2, 3):
        dom.stage(k, dt)
        orc.stage(k, dt)
        _same(dom, orc, (case, k))


@pytest.mark.parametrize("case", ["plume0", "plume1", "lRem1", "lShr"])
def test_bcon(case):
    print(f"ext bcon {case}", flush=True)
    desc, dom, orc = _pair((16, 12, 9), case, seed=5)
    dom.bcon()
    orc.bcon()
    _same(dom, orc, case)


def test_fixed_dt_odd_sizes_lrem_diffusion():
    print("ext fixed dt odd sizes", flush=True)
    cfg = _cfg((33, 17, 8), "lRem1")
    par = cfg.par.copy()
    par[55], par[56] = 0.2, 0.3
    cfg = dataclasses.replace(cfg, iD=1, par=par)
    desc = Descriptor(cfg)
    st = stratified_state(desc, seed=19)
    dom = Domain(desc)
    dom.set_state(st)
    orc = om.Oracle(desc)
    orc.set_state(st)
    orc.communicate()
    for _ in range(3):
        orc.step(fixed_dt=True, dt=2e-4)
    dom.step(3, dt=2e-4)
    _same(dom, orc, "lRem+iD fixed dt")


def test_ensemble_refuses_time_dependent_branches():
    import ctypes as C
    desc = Descriptor(_cfg((16, 12, 8), "heatT"))
    arr = (C.c_void_p * 1)(C.cast(desc.ptr(), C.c_void_p))
    h = C.c_void_p()
    rc = L.lib().mhd3d_ensemble_create(arr, 1, C.byref(h))
    assert rc == L.EUNSUPPORTED and not h.value


@pytest.mark.parametrize("flag", ["lRem", "lShr", "lMag"])
def test_from_reference_static(flag):
    """A 2 x 3 rank run's Static() (product host code) with lRem, lShr or lMag
    (AMPB layer), advanced 5 CFL steps on the GPU, against the oracle from
    the oracle's Static(): bit for bit."""
    from known_answers import default_config, default_state
    from mhd3d.initial import reference_static_state
    print(f"ext static {flag}", flush=True)
    cfg = dataclasses.replace(default_config(), nPx=16, nPy=12, nPz=18, nPz_global=20, nPy_global=12,
                              izCon=1, iTCon=1)
    par = cfg.par.copy()
    par[18] = 1.0
    par[11] = 0.0     # TP: with iTCon 1 a nonzero TP selects the plume wall (HH = 0 here)
    if flag == "lMag":
        par[6], par[7] = 300.0, 1.0                 # RM, BETA
        par[20], par[21], par[22] = 0.05, 0.3, 0.5  # AMPB, BFH, BZP
    cfg = dataclasses.replace(cfg, par=par, **{flag: True})
    got, prof = reference_static_state(cfg, 20, 2, 3, with_profiles=True)
    ref, rprof = default_state(cfg, om, NP=20, NPE_Y=2, NPE_Z=3, with_profiles=True)
    if flag == "lRem":
        cfg = dataclasses.replace(cfg, TU=prof["TU"], DZTU=prof["DZTU"], DZTB=prof["DZTB"],
                                  kappa=(prof["RKAPA"], prof["DKAPA"]))
    desc = Descriptor(cfg)
    dom = Domain(desc)
    dom.set_state(got)
    orc = om.Oracle(desc)
    orc.set_state(ref)
    orc.communicate()
    dts = [orc.step() for _ in range(5)]
    assert dom.step(5) == dts[-1]
    _same(dom, orc, flag)
