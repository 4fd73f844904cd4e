"""The RLAX heat-flux relaxation on the GPU against the oracle, bit for bit.

3dmhdsub.h:322-442 for a reference run on relax_nPEz z ranks: the mean
convective flux FCONM (one ordered plane sum per plane), the radiative flux
FRADM, the per-rank CORRECT scans, the TFAC update and the FT term, with
iTCon 0 / izCon 1 and iTCon 1 / izCon 0, hydro and MHD, with lRem, lShr and
iD 1;
10 CFL steps from a restart (TIMT past TSTART), single stages and fluxes()
through the C ABI; the statics TFAC, TTM_MAX/MIN and ICALL equal the
oracle's.  The exit points: a restart 4.5 DT before TIMT = 25 stops with
MHD3D_EEXIT in the sixth step's first fluxes() call (the state is that
after five steps), and -- with the call count shifted by mhd3d_fluxes --
in the second stage (the state is then the stage-1 output); other walls
stop as soon as the relaxation starts; set_state clears an exit and the
domain runs on with the statics it has.  Deep grids (96 planes on 8
ranks, 2800 planes on 4) take the relaxation pass's work arrays from LDS
and from global memory.  FP64 form: bit identity, 1e-12 stated as the
tolerance.
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
DT = 2.0 ** -10


def _kappa(nz):
    zs = np.linspace(0.0, 1.0, nz)
    return 1.0 + 0.5 * zs ** 2, 1.0 * zs


def _cfg(n, case, TIMI=3.7, TSTART=1.0, P=2):
    lMag = "mhd" in case
    cfg = stratified_config(*n, lMag=lMag, lRot=lMag, Ra=3e4)
    par = cfg.par.copy()
    par[57], par[58], par[59] = TSTART, 0.5, 0.2
    kw = dict(iTCon=0, izCon=1)
    if "it1" in case:
        kw = dict(iTCon=1, izCon=0)
    if "oth" in case:
        kw = dict(iTCon=4, izCon=1)
    if "rem" in case:
        kw.update(lRem=True, TU=1.0, DZTU=-0.02, DZTB=0.03, kappa=_kappa(n[2] + 2))
    if "shr" in case:
        kw.update(lShr=True)
    if "id" in case:                             # with density diffusion: RO and TT means
        kw.update(iD=1)
        par[55], par[56] = 0.5, 0.4
    return dataclasses.replace(cfg, par=par, TIMI=TIMI, relax_nPEz=P, **kw)


def _pair(cfg, seed=7):
    desc = Descriptor(cfg)
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
        assert bit_identical_fraction(g, c) == 1.0, (what, name, relerr(g, c))
        assert relerr(g, c) <= TOL


def _same_statics(dom, orc, what):
    g, c = dom.relax_state(), orc.relax_state()
    for k in ("TFAC", "TTM_MAX", "TTM_MIN", "ICALL"):
        assert g[k] == c[k], (what, k, g[k], c[k])
    assert bool(g["exited"]) == bool(c["exited"]), what


CASES = [("it0", 2), ("it0", 4), ("it1", 2), ("it0_mhd", 3), ("it1_mhd", 2), ("it0_rem", 2),
         ("it0_shr", 2), ("it1_rem_mhd", 4), ("it0_id_mhd", 2)]


@pytest.mark.parametrize("case,P", CASES)
def test_ten_steps_cfl(case, P):
    print(f"relax ten steps {case} P={P}", flush=True)
    desc, dom, orc = _pair(_cfg((24, 20, 12), case, P=P))
    dts = [orc.step() for _ in range(10)]
    assert dom.step(10) == dts[-1], case
    _same(dom, orc, case)
    _same_statics(dom, orc, case)
    assert dom.relax_state()["ICALL"] == 30
    # the relaxation acted: without it the temperature ends elsewhere
    cfg0 = _cfg((24, 20, 12), case, P=P)
    par0 = cfg0.par.copy()
    par0[59] = 0.0
    d0 = Descriptor(dataclasses.replace(cfg0, par=par0))
    o0 = om.Oracle(d0)
    o0.set_state(stratified_state(d0, seed=7))
    o0.communicate()
    for _ in range(10):
        o0.step()
    assert not np.array_equal(o0.get("TT"), orc.get("TT")), case


@pytest.mark.parametrize("n", [(128, 16, 8), (72, 10, 8), (3, 200, 6)])
def test_wide_planes(n):
    """FCONM sums of 2048, 720 and 600 values per plane (whole groups of 512,
    a group and a rest, chunks spanning many rows)."""
    print(f"relax wide planes {n}", flush=True)
    desc, dom, orc = _pair(_cfg(n, "it0", P=2), seed=11)
    dts = [orc.step() for _ in range(5)]
    assert dom.step(5) == dts[-1]
    _same(dom, orc, n)
    _same_statics(dom, orc, n)


@pytest.mark.parametrize("n,P", [((4, 4, 96), 8), ((4, 4, 2800), 4)])
def test_deep_grids(n, P):
    """Many planes and ranks: the relaxation pass's work arrays in LDS (96
    planes on 8 ranks) and, past 64 KB of them, in global memory (2800
    planes on 4 ranks)."""
    print(f"relax deep grid {n} P={P}", flush=True)
    desc, dom, orc = _pair(_cfg(n, "it0", P=P), seed=13)
    dts = [orc.step() for _ in range(3)]
    assert dom.step(3) == dts[-1]
    _same(dom, orc, (n, P))
    _same_statics(dom, orc, (n, P))


@pytest.mark.parametrize("case", ["it0_mhd", "it1_rem"])
def test_fluxes_and_stages(case):
    print(f"relax stages {case}", flush=True)
    desc, dom, orc = _pair(_cfg((20, 16, 12), case), seed=3)
    fg = dom.fluxes()
    orc.fluxes()
    for name in L.FLUXES[: dom.nf]:
        c = orc.get(name)
        assert bit_identical_fraction(fg[name], c) == 1.0, (case, name, relerr(fg[name], c))
    _same_statics(dom, orc, (case, "fluxes"))
    dt, _, _ = orc.cfl()
    for k in (1, 2, 3):
        dom.stage(k, dt)
        orc.stage(k, dt)
        _same(dom, orc, (case, k))
        _same_statics(dom, orc, (case, k))


def _exit_run(shift):
    """A restart at TIMT = 25 - 4.5 DT (every TIMT of the run exact), after
    `shift` extra fluxes() calls."""
    cfg = _cfg((24, 20, 12), "it0_mhd", TIMI=25.0 - 4.5 * DT)
    desc, dom, orc = _pair(cfg, seed=5)
    for _ in range(shift):
        dom.fluxes()
        orc.set_time(0.0, cfg.TIMI)
        orc.fluxes()
    return desc, dom, orc


def test_exit_in_the_sixth_step():
    print("relax exit at step 6 stage 1", flush=True)
    desc, dom, orc = _exit_run(0)
    for _ in range(5):
        orc.step(fixed_dt=True, dt=DT)
    with pytest.raises(L.MhdError) as ei:
        dom.step(10, dt=DT)
    assert ei.value.code == L.EEXIT
    assert dom.clock()[0] == 5
    _same(dom, orc, "state at the exit")
    orc.stage(1, DT)
    assert orc.relax_state()["exited"] == 1
    g = dom.relax_state()
    assert g["exited"] == 1 and g["ICALL"] == 16
    with pytest.raises(L.MhdError) as ei:       # the domain stays stopped
        dom.step(1, dt=DT)
    assert ei.value.code == L.EEXIT and dom.clock()[0] == 5


def test_exit_in_a_second_stage():
    print("relax exit at step 6 stage 2", flush=True)
    desc, dom, orc = _exit_run(2)                # ICALL % 3 == 0 falls on stage 2
    for _ in range(5):
        orc.step(fixed_dt=True, dt=DT)
    orc.stage(1, DT)                             # the state the exiting call reads
    with pytest.raises(L.MhdError) as ei:
        dom.step(10, dt=DT)
    assert ei.value.code == L.EEXIT and dom.clock()[0] == 5
    _same(dom, orc, "stage-2 input at the exit")
    orc.stage(2, DT)
    assert orc.relax_state()["exited"] == 1
    assert dom.relax_state()["ICALL"] == 2 + 5 * 3 + 2


def test_fluxes_call_at_the_exit_point():
    """fluxes() of a fresh run (TIMT = 0): complete fluxes, then the exit."""
    print("relax fluxes exit", flush=True)
    desc, dom, orc = _pair(_cfg((16, 12, 12), "it0", TIMI=0.0, TSTART=0.0), seed=9)
    with pytest.raises(L.MhdError) as ei:
        dom.fluxes()
    assert ei.value.code == L.EEXIT
    orc.set_time(0.0, 0.0)
    orc.fluxes()
    assert orc.relax_state()["exited"] == 1 and dom.relax_state()["exited"] == 1


def test_set_state_after_an_exit_runs_again():
    """set_state clears the exit record and keeps the other statics: after a
    fresh run's first fluxes() call exits, the next call (ICALL 1, no exit
    point) returns the reference's second call bit for bit."""
    print("relax set_state after an exit", flush=True)
    desc, dom, orc = _pair(_cfg((16, 12, 12), "it0", TIMI=0.0, TSTART=0.0), seed=9)
    st = dom.state()
    with pytest.raises(L.MhdError) as ei:
        dom.fluxes()
    assert ei.value.code == L.EEXIT and dom.relax_state()["exited"] == 1
    orc.set_time(0.0, 0.0)
    orc.fluxes()
    dom.set_state(st)
    assert dom.relax_state()["exited"] == 0
    fg = dom.fluxes()
    orc.fluxes()
    for name in L.FLUXES[: dom.nf]:
        c = orc.get(name)
        assert bit_identical_fraction(fg[name], c) == 1.0, (name, relerr(fg[name], c))
    g, c = dom.relax_state(), orc.relax_state()
    assert g["ICALL"] == c["ICALL"] == 2 and g["TFAC"] == c["TFAC"]


def test_other_walls():
    print("relax other walls", flush=True)
    # before TSTART: runs, the correction zero (FT + 0)
    desc, dom, orc = _pair(_cfg((24, 20, 12), "oth", TIMI=3.7, TSTART=50.0), seed=2)
    dts = [orc.step() for _ in range(5)]
    assert dom.step(5) == dts[-1]
    _same(dom, orc, "before TSTART")
    _same_statics(dom, orc, "before TSTART")
    # past TSTART: "Relax: check boundary temperature" in the first call
    desc, dom, orc = _pair(_cfg((24, 20, 12), "oth", TIMI=3.7, TSTART=1.0), seed=2)
    st0 = dom.state()
    with pytest.raises(L.MhdError) as ei:
        dom.step(3)
    assert ei.value.code == L.EEXIT and dom.clock()[0] == 0
    assert dom.relax_state()["exited"] == 1
    for k, v in dom.state().items():
        assert np.array_equal(v.view(np.int64), st0[k].view(np.int64)), k
