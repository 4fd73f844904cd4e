"""Ensemble members and served runs that take the off-default branches, in
the FP64 and the mixed form: iD 1/2, lRem, lShr (with and without lRem), the
iTCon 2 heat loss and the TP plume with TC = 0, and RLAX (iTCon 0 / izCon 1
and 1 / 0, one member reaching the reference's exit point), next to default
members.  Bar: every member bit-identical to the same domain advanced alone
-- state, clock, DT and the RLAX statics -- eagerly and from captured
hipGraphs, under the reference driver's DT guard; a served run the same as
alone (the exit point reported as MHD3D_REQ_EXITED).  The ensemble
autotuner leaves the RLAX statics as they were (later steps bit-identical to
an ensemble never tuned, a member reaching its exit point during a trial
included).  The heat loss and the plume with TC != 0, whose factors of TIMT
the host forms before every step, are refused.
"""
import dataclasses

import numpy as np
import pytest

import test_gpu_ext as tx
import test_gpu_relax as tr
from mhd3d import _lib as L
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.ensemble import Ensemble
from mhd3d.initial import stratified_config, stratified_state
from mhd3d.server import Server

pytestmark = pytest.mark.gpu


def _cfgs(precision=0):
    out = [tx._cfg((24, 20, 12), c) for c in ("iD1", "iD2_mhd", "lRem1", "lShr_mhd", "lShrRem", "heat0_mhd",
                                                "plume1")]
    out.append(tr._cfg((24, 20, 12), "it0_mhd", P=3))
    out.append(tr._cfg((20, 16, 12), "it1_rem", P=2))
    out.append(tr._cfg((24, 20, 12), "it0", TIMI=25.0 - 1e-9))      # exits in its second step
    out.append(tx._cfg((17, 9, 11), "lRem1"))                          # partial tiles
    out.append(tx._cfg((1, 40, 10), "iD1"))                            # one column in x
    out.append(stratified_config(32, 24, 10, Ra=5e4))
    out.append(stratified_config(28, 20, 12, lMag=True, lRot=True, Ra=1e5))
    return [dataclasses.replace(c, precision=precision) for c in out]


def _run(obj, n):
    try:
        obj.step(n)
    except L.MhdError as err:
        if err.code not in (L.EEXIT, L.EHALTED, L.ENONFINITE):
            raise


def _same(a, b):
    return all(np.array_equal(a[k].view(np.int64), b[k].view(np.int64)) for k in a)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("graph", [False, True])
def test_branch_members_bit_identical_to_alone(graph, precision):
    print(f"ensemble branches graph={graph} precision={precision}", flush=True)
    cfgs = _cfgs(precision)
    descs = [Descriptor(c) for c in cfgs]
    states = [stratified_state(d, seed=40 + i) for i, d in enumerate(descs)]
    ens = Ensemble(descs)
    for i, st in enumerate(states):
        ens.set_state(i, st)
    ens.set_dt_guard()
    ens.graph(graph)
    _run(ens, 5)
    _run(ens, 4)
    exits = 0
    for i, (d, st) in enumerate(zip(descs, states)):
        dom = Domain(d)
        dom.set_state(st)
        dom.set_dt_guard()
        _run(dom, 5)
        _run(dom, 4)
        assert _same(ens.state(i), dom.state()), (graph, i)
        n, timt, timc = dom.clock()
        n_e, timt_e, timc_e, _ = ens.clock(i)
        assert (n_e, timt_e, timc_e) == (n, timt, timc), (i, n_e, n)
        if d.cfg.par[59] > 0:
            g = dom.relax_state()
            assert g["ICALL"] > 0, i
            exits += g["exited"]
        dom.close()
    assert exits == 1
    ens.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_time_dependent_branches_refused(precision):
    desc = Descriptor(dataclasses.replace(tx._cfg((16, 12, 8), "heatT"), precision=precision))
    with pytest.raises(L.MhdError) as ei:
        Ensemble([desc])
    assert ei.value.code == L.EUNSUPPORTED
    srv = Server(precision, 1 << 30, 4)
    with pytest.raises(L.MhdError) as ei:
        srv.submit(desc, 5, 1)
    assert ei.value.code == L.EUNSUPPORTED
    srv.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_served_branch_runs_identical_to_alone(precision):
    print(f"server branches precision={precision}", flush=True)
    cfgs = _cfgs(precision)
    steps = [6 + 3 * i for i in range(len(cfgs))]
    srv = Server(precision, 1 << 30, 4)
    ids = []
    for i, c in enumerate(cfgs):
        ids.append(srv.submit(Descriptor(c), steps[i], 500 + i))
    status = {}
    for _ in range(200):
        srv.tick()
        for rid in ids:
            if rid in status:
                continue
            info = srv.poll(rid)
            if info["status"] in (L.REQ_DONE, L.REQ_FAILED, L.REQ_NONFINITE, L.REQ_EXITED):
                status[rid] = info["status"]
        if len(status) == len(ids):
            break
    assert len(status) == len(ids)
    for i, (c, rid) in enumerate(zip(cfgs, ids)):
        got = srv.result(rid, fetch=status[rid] == L.REQ_DONE)
        dom = Domain(Descriptor(c))
        dom.init_stratified(seed=500 + i)
        dom.set_dt_guard()
        code = L.OK
        try:
            dom.step(steps[i])
        except L.MhdError as err:
            code = err.code
        want = {L.OK: L.REQ_DONE, L.EHALTED: L.REQ_FAILED, L.ENONFINITE: L.REQ_NONFINITE,
                L.EEXIT: L.REQ_EXITED}[code]
        assert status[rid] == want, (i, status[rid], code)
        if got is not None:
            assert _same(got, dom.state()), i
        dom.close()
    srv.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_keeps_the_relaxation_statics(precision):
    """Two RLAX members (one reaching its exit point in its second step,
    inside the tuner's trials) and a default one: tuned after two steps, then
    five more, against the same ensemble never tuned."""
    print(f"ensemble autotune rlax precision={precision}", flush=True)
    cfgs = [tr._cfg((24, 20, 12), "it0_mhd", P=3), tr._cfg((24, 20, 12), "it0", TIMI=25.0 - 1e-9),
            stratified_config(32, 24, 10, Ra=5e4)]
    descs = [Descriptor(dataclasses.replace(c, precision=precision)) for c in cfgs]
    states = [stratified_state(d, seed=70 + i) for i, d in enumerate(descs)]
    runs = []
    for tune in (True, False):
        ens = Ensemble(descs)
        for i, st in enumerate(states):
            ens.set_state(i, st)
        ens.step(1)
        if tune:
            _, trials = ens.autotune(steps=2)
            assert len(trials) == 32 and all(t["ms_per_step"] > 0 for t in trials)
        _run(ens, 5)
        runs.append(([ens.state(i) for i in range(len(descs))], [ens.clock(i) for i in range(len(descs))]))
        ens.close()
    (st_t, ck_t), (st_u, ck_u) = runs
    for i in range(len(descs)):
        assert _same(st_t[i], st_u[i]), i
        assert ck_t[i] == ck_u[i], (i, ck_t[i], ck_u[i])
    assert ck_u[1][0] < 6                       # the exiting member stopped early
