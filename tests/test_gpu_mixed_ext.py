"""The mixed form with the off-default branches, on the GPU, against the
oracle (FP64).

With a branch switched on, the mixed form keeps its storage -- FP32 copies of
the stage values (deviations from the FP64 background of each plane), FP64
RK registers -- and evaluates the fluxes in FP64 from those copies with the
background added back, so the branches run through the FP64 code; the plane
means inside a step are fixed-order trees of partial sums instead of the
reference's one chain (the same doubles alone and in an ensemble).  Bar:
after 10 CFL steps every field within 1e-5 of the oracle
(max|g - c| / max|c| over the interior incl. walls; the FP32 rounding of the
stage values gives ~1e-7), for iD 1 and 2, lRem, lShr (hydro and MHD), lShr
with lRem (also at one column in x and at odd sizes with partial tiles),
the iTCon 2 heat loss with its time factor, the TP plume, and
RLAX (both wall settings, with lRem and MHD); RLAX's statics follow the
oracle's (the call count exactly, TFAC to 1e-12).  At an RLAX exit point in a
stage 1 or a stage 2 the run stops with MHD3D_EEXIT and keeps the state of
the step's start, its one FP64 register (the stage values live only in the
FP32 copies): bit-identical to the same run advanced up to that step.  Under lShr the walls
drive a shear flow (|RV| ~ 10, stored in FP32 as it is: only RO, TT and B
are stored as deviations from a background) next to |RU| ~ 1e-3, so RU's
error relative to its own maximum is bounded by 5e-5 there (measured
1.2e-5).
"""
import dataclasses

import numpy as np
import pytest

import test_gpu_ext as tx
import test_gpu_relax as tr
from gpu_common import relerr
from mhd3d import _lib as L
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.initial import stratified_state
from oracle import oracle as om

pytestmark = pytest.mark.gpu

TOL = 1e-5
TOL_SHEAR = 5e-5

CASES = ["iD1", "iD2_mhd", "lRem1", "lShr", "lShr_mhd", "lShrRem", "heatT", "plume0", "iDrem_mhd"]
RLAX_CASES = [("it0", 2), ("it1_rem_mhd", 4)]


def _run(cfg, seed, what, tol=TOL):
    d64 = Descriptor(cfg)
    dmx = Descriptor(dataclasses.replace(cfg, precision=1))
    st = stratified_state(d64, seed=seed)
    orc = om.Oracle(d64)
    orc.set_state(st)
    orc.communicate()
    for _ in range(10):
        orc.step()
    dom = Domain(dmx)
    dom.set_state(st)
    dom.step(10)
    worst = 0.0
    for name, g in dom.state().items():
        e = relerr(g, orc.get(name))
        worst = max(worst, e)
        assert e <= tol, (what, name, e)
    print(f"  {what}: worst field error {worst:.2e}", flush=True)
    return dom, orc


@pytest.mark.parametrize("case", CASES)
def test_branch_in_mixed_form(case):
    print(f"mixed branch {case}", flush=True)
    dom, orc = _run(tx._cfg((24, 20, 12), case), 7, case, TOL_SHEAR if case.startswith("lShr") else TOL)
    dom.close()
    orc.close()


@pytest.mark.parametrize("case,n", [("iD1", (1, 40, 10)), ("lRem1", (17, 9, 11)), ("lShr_mhd", (33, 7, 9))])
def test_branch_in_mixed_form_odd_sizes(case, n):
    """One column in x, and odd sizes that leave partial tiles in x and y."""
    print(f"mixed branch {case} {n}", flush=True)
    dom, orc = _run(tx._cfg(n, case), 13, (case, n), TOL_SHEAR if case.startswith("lShr") else TOL)
    dom.close()
    orc.close()


@pytest.mark.parametrize("case,P", RLAX_CASES)
def test_rlax_in_mixed_form(case, P):
    print(f"mixed rlax {case} P={P}", flush=True)
    dom, orc = _run(tr._cfg((24, 20, 12), case, P=P), 7, (case, P))
    g, c = dom.relax_state(), orc.relax_state()
    assert g["ICALL"] == c["ICALL"] == 30
    assert abs(g["TFAC"] - c["TFAC"]) <= 1e-12 * abs(c["TFAC"])
    dom.close()
    orc.close()


@pytest.mark.parametrize("shift", [0, 2])
def test_rlax_exit_in_mixed_form(shift):
    """shift 0: the exit falls on the sixth step's stage 1; shift 2 (two
    fluxes() calls first): on its stage 2."""
    print(f"mixed rlax exit shift={shift}", flush=True)
    cfg = dataclasses.replace(tr._cfg((24, 20, 12), "it0_mhd", TIMI=25.0 - 4.5 * tr.DT), precision=1)
    desc = Descriptor(cfg)
    st = stratified_state(desc, seed=5)
    dom, ref = Domain(desc), Domain(desc)
    for d in (dom, ref):
        d.set_state(st)
        for _ in range(shift):
            d.fluxes()
    ref.step(5, dt=tr.DT)
    with pytest.raises(L.MhdError) as ei:
        dom.step(10, dt=tr.DT)
    assert ei.value.code == L.EEXIT and dom.clock()[0] == 5
    g = dom.relax_state()
    assert g["exited"] == 1 and g["ICALL"] == shift + 5 * 3 + (2 if shift else 1)
    want = ref.state()
    for k, v in dom.state().items():
        assert np.array_equal(v.view(np.int64), want[k].view(np.int64)), (shift, k)
    dom.close()
    ref.close()
