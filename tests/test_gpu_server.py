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


@pytest.mark.parametrize("nsteps", [4, 5])
def test_dt_guard_halts_after_the_step_like_the_driver(nsteps):
    """3dmhd.cpp:292: after a step with DT < threshold the run halts; the
    state is that of the step which tripped the guard, whatever the parity
    of the number of steps asked for after it."""
    from mhd3d import _lib as L
    d = Descriptor(stratified_config(32, 24, 12, lMag=True, lRot=True))
    ref = Domain(d)
    ref.init_stratified(seed=5)
    ref.step(1)
    dom = Domain(d)
    dom.init_stratified(seed=5)
    dom.set_dt_guard(1.0)           # every DT is below 1: halts after step 1
    with pytest.raises(L.MhdError) as ei:
        dom.step(nsteps)
    assert ei.value.code == L.EHALTED
    assert dom.clock()[0] == 1
    assert _same(dom.state(), ref.state())
    with pytest.raises(L.MhdError):  # a halted run stays where it is
        dom.step(3)
    assert _same(dom.state(), ref.state())
    dom.init_stratified(seed=5)     # a new state runs again
    dom.set_dt_guard(0.0)
    dom.step(2)
    assert dom.clock()[0] == 3


def test_resumed_run_takes_a_fresh_tstep():
    """Lowering the guard of a halted run resumes it; its next DT is the
    TSTEP of its state, not the DT of the step that tripped the guard."""
    from mhd3d import _lib as L
    d = Descriptor(stratified_config(32, 24, 12, lMag=True, lRot=True))
    dom = Domain(d)
    dom.init_stratified(seed=6)
    dom.set_dt_guard(1.0)
    with pytest.raises(L.MhdError):
        dom.step(2)
    tripped = dom.state()
    dom.set_dt_guard(0.0)
    dt_cfl, _, _ = dom.cfl()
    dt = dom.step(1)
    ref = Domain(d)
    ref.set_state(tripped)
    dt_ref = ref.step(1)
    assert dt == dt_cfl == dt_ref
    assert _same(dom.state(), ref.state())


def test_dt_guard_in_an_ensemble_stops_one_member():
    """Only member 0 gets a guard every DT trips: it halts after its first
    step and stays there; member 1 advances as it does alone."""
    from mhd3d.ensemble import Ensemble
    from mhd3d.initial import stratified_state
    descs = [Descriptor(stratified_config(24, 16, 9)), Descriptor(stratified_config(32, 24, 12, lMag=True))]
    states = [stratified_state(d, seed=9 + k) for k, d in enumerate(descs)]
    ens = Ensemble(descs)
    for i, st in enumerate(states):
        ens.set_state(i, st)
    ens.set_dt_guard(1.0, member=0)
    ens.step(2)
    ens.step(3)
    assert ens.halted(0) and not ens.halted(1)
    assert ens.clock(0)[0] == 1 and ens.clock(1)[0] == 5
    for i, n in ((0, 1), (1, 5)):
        alone = Domain(descs[i])
        alone.set_state(states[i])
        alone.step(n)
        assert _same(ens.state(i), alone.state()), i
    # resumed, the halted member advances again in step with itself alone
    ens.set_dt_guard(0.0, member=0)
    ens.step(3)
    assert ens.clock(0)[0] == 4 and ens.clock(1)[0] == 8
    for i, n in ((0, 4), (1, 8)):
        alone = Domain(descs[i])
        alone.set_state(states[i])
        alone.step(n)
        assert _same(ens.state(i), alone.state()), i


def test_ensemble_guard_survives_set_state():
    """set_state resets a member's clock and keeps its DT guard: a guard set
    before the states are loaded still halts member 0 after its first
    step."""
    from mhd3d.ensemble import Ensemble
    from mhd3d.initial import stratified_state
    descs = [Descriptor(stratified_config(24, 16, 9)), Descriptor(stratified_config(32, 24, 12, lMag=True))]
    states = [stratified_state(d, seed=9 + k) for k, d in enumerate(descs)]
    ens = Ensemble(descs)
    ens.set_dt_guard(1.0, member=0)
    for i, st in enumerate(states):
        ens.set_state(i, st)
    ens.step(3)
    assert ens.halted(0) and not ens.halted(1)
    assert ens.clock(0)[0] == 1 and ens.clock(1)[0] == 3
    ens.close()


def _spec(i, g, steps, lMag=False):
    return RunSpec(index=i, nPx=g[0], nPy=g[1], nPz=g[2], Ra=1e5, Pr=1.0, Pm=3.0, lMag=lMag,
                   lRot=True, seed=700 + i, steps=steps, arrival=0.0)


def test_result_fetched_after_later_ticks():
    """A run that finishes inside a quantum stops on the device at its own
    step count; fetched only after the other runs have ticked on by an odd
    number of steps, it is still the run alone."""
    from mhd3d import _lib as L
    specs = [_spec(0, (24, 16, 9), 2), _spec(1, (32, 24, 12), 21, lMag=True), _spec(2, (16, 16, 8), 14)]
    srv = Server(0, hbm_budget_bytes=1 << 30, quantum_steps=5)
    descs = [Descriptor(s.config(0)) for s in specs]
    ids = [srv.submit(d, s.steps, s.seed) for d, s in zip(descs, specs)]
    srv.tick()
    assert srv.poll(ids[0])["status"] == L.REQ_DONE and srv.poll(ids[0])["steps_done"] == 2
    while srv.tick()[0] > 0:
        pass
    for rid, s in zip(ids, specs):
        info = srv.poll(rid)
        assert info["status"] == L.REQ_DONE and info["steps_done"] == s.steps, (s.index, info)
        assert _same(srv.result(rid), _alone(s, 0)), s.index
    srv.close()


def test_nonfinite_run_is_evicted():
    """A run whose state holds NaN is evicted with its own status after its
    first tick, while the other run finishes as alone."""
    from mhd3d import _lib as L
    from mhd3d.initial import stratified_profiles
    import ctypes as C
    good = _spec(0, (24, 16, 9), 12)
    bad = _spec(1, (16, 16, 8), 12)
    srv = Server(0, hbm_budget_bytes=1 << 30, quantum_steps=4)
    d_good = Descriptor(good.config(0))
    d_bad = Descriptor(bad.config(0))
    id_good = srv.submit(d_good, good.steps, good.seed)
    T, rho = stratified_profiles(d_bad)
    T = np.ascontiguousarray(T)
    rho = np.ascontiguousarray(rho)
    T[4] = np.nan
    rid = C.c_int64()
    dp = C.POINTER(C.c_double)
    srv._check(L.lib().mhd3d_server_submit(srv.h, C.cast(d_bad.ptr(), C.c_void_p), T.ctypes.data_as(dp),
                                           rho.ctypes.data_as(dp), int(bad.seed), 1e-4, 1e-3, 1e-5,
                                           int(bad.steps), C.byref(rid)))
    srv._descs[rid.value] = d_bad
    srv.tick()
    assert srv.poll(rid.value)["status"] == L.REQ_NONFINITE
    srv.result(rid.value, fetch=False)
    while srv.tick()[0] > 0:
        pass
    assert srv.poll(id_good)["status"] == L.REQ_DONE
    assert _same(srv.result(id_good), _alone(good, 0))
    srv.close()


@pytest.mark.parametrize("fused", [True, False])
def test_nonfinite_flag_follows_the_state(fused):
    """MHD3D_ENONFINITE reports NaN/Inf in the state, from the TSTEP partials
    of the last step (CFL steps) or a scan (fixed DT); a finite state whose
    CFL limits are NaN (a negative temperature under the square root) steps
    on as in the reference, whose minima ignore NaN."""
    from mhd3d import _lib as L
    from mhd3d.initial import stratified_state
    d = Descriptor(stratified_config(24, 16, 9, lMag=True))
    st = stratified_state(d, seed=21)
    dom = Domain(d)
    neg = {k: v.copy() for k, v in st.items()}
    neg["TT"][5, 6, 7] = -0.5
    dom.set_state(neg)
    dt = None if fused else 1e-4
    dom.step(2, dt=dt)
    assert all(np.isfinite(v).all() for v in dom.state().values())
    bad = {k: v.copy() for k, v in st.items()}
    bad["TT"][5, 6, 7] = np.inf
    dom.set_state(bad)
    with pytest.raises(L.MhdError) as ei:
        dom.step(1, dt=dt)
    assert ei.value.code == L.ENONFINITE
