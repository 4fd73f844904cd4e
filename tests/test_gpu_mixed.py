"""Mixed form: FP32 flux stencils from FP32 copies of the stage values, FP64
RK3 registers (U, Z), FP64 CFL and FP64 diagnostics.

Bars (error = max|g-c|/max|c| per field over the interior incl. walls):
  * 100 steps (DT from the CFL each step) against the FP64 form on a
    128x128x64 stratified run: <= 1e-5 per field; total mass and total energy
    within 1e-9 relative; the DT of the last step within 1e-6 relative;
  * 10 steps against the oracle: <= 1e-6 per field;
  * the FP64 state handed in comes back unchanged (set/get keep FP64).
"""
import numpy as np
import pytest

from gpu_common import CASES, make_pair, relerr
from mhd3d import _lib as L
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.initial import stratified_config, stratified_state

pytestmark = pytest.mark.gpu

TOL_100 = 1e-5
TOL_ORACLE_10 = 1e-6
TOL_INTEGRALS = 1e-9


def _pair_mixed(n, case, seed, **kw):
    """FP64 and mixed domains on the same seeded stratified state."""
    out = []
    st = None
    for prec in (0, 1):
        cfg = stratified_config(*n, **CASES[case], precision=prec, **kw)
        desc = Descriptor(cfg)
        if st is None:
            st = stratified_state(desc, seed=seed)
        dom = Domain(desc)
        dom.set_state(st)
        out.append(dom)
    return out[0], out[1], st


@pytest.mark.parametrize("case", ["hydro", "mhd_rot"])
def test_mixed_vs_fp64_100_steps(case):
    d64, dmx, _ = _pair_mixed((128, 128, 64), case, seed=23)
    dt64 = d64.step(100)
    dtmx = dmx.step(100)
    assert abs(dtmx - dt64) <= 1e-6 * abs(dt64), (dt64, dtmx)
    s64 = d64.state()
    smx = dmx.state()
    errs = {k: relerr(smx[k], s64[k]) for k in s64}
    print(case, "per-field error after 100 steps:", errs)
    for k, e in errs.items():
        assert e <= TOL_100, (case, k, e)
    g64 = d64.diagnostics()
    gmx = dmx.diagnostics()
    print(case, "mass/energy/max|divB| FP64:", g64, "mixed:", gmx)
    for k in ("mass", "energy"):
        assert abs(gmx[k] - g64[k]) <= TOL_INTEGRALS * abs(g64[k]), (k, g64[k], gmx[k])


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_vs_oracle_10_steps(case):
    desc, _, orc, st = make_pair((32, 24, 16), case, seed=11)
    cfg = stratified_config(32, 24, 16, **CASES[case], precision=1)
    dmx = Domain(Descriptor(cfg))
    dmx.set_state(st)
    for _ in range(10):
        orc.step()
    dmx.step(10)
    for name, g in dmx.state().items():
        e = relerr(g, orc.get(name))
        assert e <= TOL_ORACLE_10, (case, name, e)


@pytest.mark.parametrize("case,n", [("hydro", (33, 17, 7)), ("mhd_rot", (33, 17, 7)), ("hydro", (1, 24, 12)),
                                    ("mhd_rot", (1, 20, 9))])
def test_mixed_fixed_dt_odd_sizes(case, n):
    """Odd sizes and one column in x (nPx = 1: every x difference vanishes)."""
    print(f"mixed odd sizes {case} {n}", flush=True)
    desc, _, orc, st = make_pair(n, case, seed=19)
    cfg = stratified_config(*n, **CASES[case], precision=1)
    dmx = Domain(Descriptor(cfg))
    dmx.set_state(st)
    dt = 1e-4
    for _ in range(3):
        orc.step(fixed_dt=True, dt=dt)
    dmx.step(3, dt=dt)
    for name, g in dmx.state().items():
        e = relerr(g, orc.get(name))
        assert e <= TOL_ORACLE_10, (case, name, e)
    n, _, timc = dmx.clock()
    assert n == 3 and timc == 3 * dt


def test_mixed_state_roundtrip_and_stage_refused():
    cfg = stratified_config(20, 16, 9, lMag=True, precision=1)
    desc = Descriptor(cfg)
    st = stratified_state(desc, seed=4)
    dom = Domain(desc)
    dom.set_state(st)
    back = dom.state()
    for k, v in back.items():
        a = np.asarray(st[k])[1:-1, 1:-1, 1:-1]
        assert np.array_equal(v[1:-1, 1:-1, 1:-1], a), k
    with pytest.raises(L.MhdError) as ei:
        dom.stage(1, 1e-4)
    assert ei.value.code == L.EUNSUPPORTED


@pytest.mark.parametrize("tune", ["1,0", "1,1", "1,2", "1,3", "2,0", "2,1", "2,2", "2,3"])
@pytest.mark.parametrize("case", ["hydro", "mhd_rot"])
def test_mixed_schedule_variants(case, tune, monkeypatch):
    """Every streaming-kernel variant of the mixed form meets the oracle bar."""
    desc, _, orc, st = make_pair((40, 24, 12), case, seed=29)
    env = "MHD3D_TUNE_MIXED_" + ("MHD" if "mhd" in case else "HYDRO")
    monkeypatch.setenv(env, tune)
    dmx = Domain(Descriptor(stratified_config(40, 24, 12, **CASES[case], precision=1)))
    dmx.set_state(st)
    for _ in range(3):
        orc.step()
    dmx.step(3)
    for name, g in dmx.state().items():
        e = relerr(g, orc.get(name))
        assert e <= TOL_ORACLE_10, (case, tune, name, e)
