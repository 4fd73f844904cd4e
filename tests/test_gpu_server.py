"""Device initial state and the request server.  Bars: the device initializer
equals its host copy bit for bit; every served run (queued behind an HBM
budget, advanced in a changing batch) is bit-identical to the same run
advanced alone; the budget is never exceeded."""
import numpy as np
import pytest

from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.initial import stratified_config, stratified_state_hashed
from mhd3d.server import Server, serve_trace
from mhd3d.workloads import RunSpec

pytestmark = pytest.mark.gpu


def _same(a, b):
    return all(np.array_equal(a[k].view(np.int64), b[k].view(np.int64)) for k in a)


@pytest.mark.parametrize("lMag", [False, True])
def test_device_init_matches_host_copy(lMag):
    d = Descriptor(stratified_config(40, 24, 12, lMag=lMag, lRot=True))
    dom = Domain(d)
    dom.init_stratified(seed=77)
    got = dom.state()
    ref = stratified_state_hashed(d, seed=77)
    inner = (slice(None), slice(1, -1), slice(1, -1))
    for k in got:
        assert np.array_equal(got[k][inner], ref[k][inner]), k


def _specs():
    grids = [(24, 16, 9), (40, 24, 12), (33, 17, 7), (16, 32, 10), (64, 40, 20), (32, 32, 16)]
    out = []
    for i, g in enumerate(grids):
        # magnetic Reynolds numbers RM = Pm*RE >= 100: at lower RM the wall
        # planes' one-sided magnetic diffusion (3dmhdsub.h:638-690) grows on
        # these coarse grids, alone as in a batch
        out.append(RunSpec(index=i, nPx=g[0], nPy=g[1], nPz=g[2], Ra=10 ** (5 + i % 2), Pr=0.5 + 0.1 * i,
                           Pm=2.0 + 0.2 * i, lMag=(i % 2 == 1), lRot=(i % 3 == 0), seed=300 + i,
                           steps=5 + 7 * i, arrival=0.0))
    return out


def _alone(spec, precision):
    dom = Domain(Descriptor(spec.config(precision)))
    dom.init_stratified(seed=spec.seed)
    dom.step(spec.steps)
    return dom.state()


@pytest.mark.parametrize("precision", [0, 1])
def test_served_runs_identical_to_alone(precision):
    specs = _specs()
    # a budget of about two runs forces queuing and a changing batch
    budget = 2 * 3 * 8 * 8 * 66 * 42 * 22 + (4 << 20)
    out = serve_trace(specs, precision=precision, hbm_budget_bytes=budget, quantum_steps=4, fetch=True)
    assert out["peak_device_bytes"] <= budget
    for i, s in enumerate(specs):
        assert _same(out["results"][i], _alone(s, precision)), (precision, i)


def test_oversized_request_refused():
    srv = Server(0, hbm_budget_bytes=1 << 20, quantum_steps=4)
    d = Descriptor(stratified_config(64, 64, 32))
    with pytest.raises(Exception):
        srv.submit(d, 10, seed=1)


def test_dt_guard_halts_after_the_step_like_the_driver():
    """3dmhd.cpp:292: after a step with DT < threshold the run halts; the
    state is that of the step which tripped the guard."""
    from mhd3d import _lib as L
    d = Descriptor(stratified_config(32, 24, 12, lMag=True, lRot=True))
    ref = Domain(d)
    ref.init_stratified(seed=5)
    ref.step(1)
    dom = Domain(d)
    dom.init_stratified(seed=5)
    dom.set_dt_guard(1.0)           # every DT is below 1: halts after step 1
    with pytest.raises(L.MhdError) as ei:
        dom.step(5)
    assert ei.value.code == L.EHALTED
    assert dom.clock()[0] == 1
    assert _same(dom.state(), ref.state())
    dom.init_stratified(seed=5)     # a new state runs again
    dom.set_dt_guard(0.0)
    dom.step(2)
    assert dom.clock()[0] == 3


def test_dt_guard_in_an_ensemble_stops_one_member():
    from mhd3d.ensemble import Ensemble
    from mhd3d.initial import stratified_state
    descs = [Descriptor(stratified_config(24, 16, 9)), Descriptor(stratified_config(32, 24, 12, lMag=True))]
    states = [stratified_state(d, seed=9 + k) for k, d in enumerate(descs)]
    ens = Ensemble(descs)
    for i, st in enumerate(states):
        ens.set_state(i, st)
    ens.set_dt_guard(0.0)
    # member 1 gets a guard every DT trips
    from mhd3d import _lib as L
    import ctypes as C
    ens.step(1)
    ens.set_dt_guard(1.0)
    ens.step(3)
    assert ens.halted(0) and ens.halted(1)
    assert ens.clock(0)[0] == 2 and ens.clock(1)[0] == 2
    for i, (d, st) in enumerate(zip(descs, states)):
        alone = Domain(d)
        alone.set_state(st)
        alone.step(2)
        assert _same(ens.state(i), alone.state()), i
