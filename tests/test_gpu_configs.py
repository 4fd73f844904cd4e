"""The workload configurations themselves on the GPU, against the oracle.

* Plain FP64 hydro 96x96x64 (tanh z, fixed wall temperatures, Ra 1e5, Pr 1),
  10 steps at the fixed DT of the initial state: after every step the GPU
  state is dumped in the reference's .dat0 record layout (3dmhdio.h:20-35)
  and compared record by record with the oracle's state in the same layout:
  max relative error <= 1e-12 per field, and every value bit-identical.
* The 64-run ensemble mix (nx = ny in {32..128}, nz in {32..96}, Ra 1e4-1e6,
  half lMag), 200 steps in one persistent launch per stage, replayed from a
  hipGraph, FP64 and mixed, every member under the reference driver's DT
  guard: every member bit-identical to itself alone after 200 steps, with
  the same step count; the smallest hydro and the smallest MHD member
  against the oracle after 10 steps (FP64 <= 1e-12, mixed <= 1e-5 per
  field).  Some members of the mix are under-resolved convection that runs
  away within 200 steps (DT collapsing below the guard, or NaN in the mixed
  form) -- in the reference's arithmetic too (the oracle shows it) -- and
  those must also be the same alone and in the ensemble, bit for bit.
* The served trace (40 runs, Poisson arrivals 50/s, grids 32^3 to
  256x256x128, 50-2000 steps), FP64: every finished run bit-identical to the
  same run alone, runs the DT guard halts halted at the same step alone, and
  the two cheapest runs (one hydro, one MHD) against the oracle.
"""
import numpy as np
import pytest

from gpu_common import bit_identical_fraction, relerr
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.dumpio import RECORD_FIELDS, dump_data
from mhd3d.ensemble import Ensemble
from mhd3d.initial import stratified_config, stratified_state, stratified_state_hashed
from mhd3d.workloads import ensemble_mix, served_trace
from oracle import oracle as om

pytestmark = pytest.mark.gpu


def _same(a, b):
    return all(np.array_equal(a[k].view(np.int64), b[k].view(np.int64)) for k in a)


def _records(path, nf, shape, idx):
    nz, ny, nx = shape
    core = (nz - 2) * (ny - 2) * (nx - 2)
    raw = np.fromfile(path, dtype="<f8", count=nf * core, offset=idx * nf * core * 8)
    return raw.reshape(nf, nz - 2, ny - 2, nx - 2)


def test_hydro_96x96x64_fixed_dt_dumps_match_per_iteration(tmp_path):
    cfg = stratified_config(96, 96, 64, Ra=1e5, Pr=1.0)
    desc = Descriptor(cfg)
    st = stratified_state(desc, seed=2024)
    dom = Domain(desc)
    dom.set_state(st)
    orc = om.Oracle(desc)
    orc.set_state(st)
    orc.communicate()
    dt, _, _ = orc.cfl()          # the fixed DT: TSTEP of the initial state
    gpu_dump = str(tmp_path / "gpu.dat0.000")
    nf = 5
    for it in range(10):
        dom.step(1, dt=dt)
        orc.step(fixed_dt=True, dt=dt)
        dump_data(dom, gpu_dump, it)
        rec = _records(gpu_dump, nf, desc.shape, it)
        for r, name in enumerate(RECORD_FIELDS[:nf]):
            c = orc.get(name)[1:-1, 1:-1, 1:-1]
            g = rec[r]
            err = float(np.abs(g - c).max() / np.abs(c).max()) if np.abs(c).max() > 0 else float(np.abs(g).max())
            assert err <= 1e-12, (it, name, err)
            assert np.array_equal(g.view(np.int64), c.view(np.int64)), (it, name)
    assert dom.clock()[0] == 10
    orc.close()


def _run(obj, n):
    """step(n), tolerating what the reference's driver tolerates or stops
    at: a non-finite state (its minima ignore NaN) and the DT guard."""
    from mhd3d import _lib as L
    try:
        obj.step(n)
    except L.MhdError as err:
        if err.code not in (L.ENONFINITE, L.EHALTED):
            raise


def _mix_members(precision, specs):
    descs = [Descriptor(s.config(precision)) for s in specs]
    states = [stratified_state_hashed(d, seed=s.seed) for d, s in zip(descs, specs)]
    return descs, states


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_mix_64_runs_200_steps(precision):
    specs = ensemble_mix(64)
    descs, states = _mix_members(precision, specs)
    ens = Ensemble(descs)
    for i, s in enumerate(states):
        ens.set_state(i, s)
    ens.set_dt_guard()
    ens.graph(True)
    _run(ens, 10)
    # the cheapest hydro and MHD members against the oracle after 10 steps
    cells = [s.cells for s in specs]
    picks = [min((i for i in range(64) if not specs[i].lMag), key=lambda i: cells[i]),
             min((i for i in range(64) if specs[i].lMag), key=lambda i: cells[i])]
    tol = 1e-12 if precision == 0 else 1e-5
    for i in picks:
        d64 = Descriptor(specs[i].config(0))
        orc = om.Oracle(d64)
        orc.set_state(states[i])
        orc.communicate()
        for _ in range(10):
            orc.step()
        got = ens.state(i)
        for name, g in got.items():
            e = relerr(g, orc.get(name))
            assert e <= tol, (precision, i, name, e)
            if precision == 0:
                assert bit_identical_fraction(g, orc.get(name)) == 1.0, (i, name)
        orc.close()
    _run(ens, 190)
    for i, (d, s) in enumerate(zip(descs, states)):
        alone = Domain(d)
        alone.set_state(s)
        alone.set_dt_guard()
        _run(alone, 10)
        _run(alone, 190)
        assert _same(ens.state(i), alone.state()), (precision, i)
        n = alone.clock()[0]
        assert ens.clock(i)[0] == n, (i, ens.clock(i)[0], n)
        assert n == 200 or ens.halted(i), i
        alone.close()
    ens.close()


def test_served_trace_40_runs():
    from mhd3d import _lib as L
    from mhd3d.server import serve_trace
    specs = served_trace(40, rate=50.0)
    out = serve_trace(specs, precision=0, quantum_steps=10, fetch=True)
    failed = set(out["failed"])
    evicted = set(out["evicted_nonfinite"])
    for i, s in enumerate(specs):
        dom = Domain(Descriptor(s.config(0)))
        dom.init_stratified(seed=s.seed)
        dom.set_dt_guard()                     # the guard every served run gets
        code = L.OK
        try:
            dom.step(s.steps)
        except L.MhdError as err:
            assert err.code in (L.EHALTED, L.ENONFINITE), err
            code = err.code
        assert (code == L.EHALTED) == (i in failed), i
        assert (code == L.ENONFINITE) == (i in evicted), i
        if code == L.OK:
            assert _same(out["results"][i], dom.state()), i
        dom.close()
    # the two cheapest runs that finished (one hydro, one MHD) against the oracle
    done = [i for i in range(len(specs)) if i not in failed and i not in evicted]
    for mag in (False, True):
        i = min((j for j in done if specs[j].lMag == mag), key=lambda j: specs[j].cells * specs[j].steps)
        desc = Descriptor(specs[i].config(0))
        st = stratified_state_hashed(desc, seed=specs[i].seed)
        orc = om.Oracle(desc)
        orc.set_state(st)
        orc.communicate()
        for _ in range(specs[i].steps):
            orc.step()
        for name, g in out["results"][i].items():
            e = relerr(g, orc.get(name))
            assert e <= 1e-12, (i, name, e)
        orc.close()
