"""Ensemble members and served runs that take the off-default branches
(FP64): iD 1/2, lRem, lShr (with and without lRem), the iTCon 2 heat loss
and the TP plume with TC = 0, and RLAX (iTCon 0 / izCon 1 and 1 / 0, one
member reaching the reference's exit point), next to default members.
Bar: every member bit-identical to the same domain advanced alone -- state,
clock, DT and the RLAX statics -- eagerly and from captured hipGraphs, under
the reference driver's DT guard; a served run the same as alone (the exit
point reported as MHD3D_REQ_EXITED).  The heat loss and the plume with
TC != 0, whose factors of TIMT the host forms before every step, are
refused.
"""

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


def _cfgs():
    out = [tx._cfg((24, 20, 12), c) for c in ("iD1", "iD2_mhd", "lRem1", "lShr_mhd", "lShrRem", "heat0_mhd",
                                                "plume1")]
    out.append(tr._cfg((24, 20, 12), "it0_mhd", P=3))
    out.append(tr._cfg((20, 16, 12), "it1_rem", P=2))
    out.append(tr._cfg((24, 20, 12), "it0", TIMI=25.0 - 1e-9))      # exits in its second step
    out.append(stratified_config(32, 24, 10, Ra=5e4))
    out.append(stratified_config(28, 20, 12, lMag=True, lRot=True, Ra=1e5))
    return out


def _run(obj, n):
    try:
        obj.step(n)
    except L.MhdError as err:
        if err.code not in (L.EEXIT, L.EHALTED, L.ENONFINITE):
            raise


def _same(a, b):
    return all(np.array_equal(a[k].view(np.int64), b[k].view(np.int64)) for k in a)


@pytest.mark.parametrize("graph", [False, True])
def test_branch_members_bit_identical_to_alone(graph):
    print(f"ensemble branches graph={graph}", flush=True)
    cfgs = _cfgs()
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


def test_time_dependent_branches_refused():
    desc = Descriptor(tx._cfg((16, 12, 8), "heatT"))
    with pytest.raises(L.MhdError) as ei:
        Ensemble([desc])
    assert ei.value.code == L.EUNSUPPORTED
    srv = Server(0, 1 << 30, 4)
    with pytest.raises(L.MhdError) as ei:
        srv.submit(desc, 5, 1)
    assert ei.value.code == L.EUNSUPPORTED
    srv.close()


def test_served_branch_runs_identical_to_alone():
    print("server branches", flush=True)
    cfgs = _cfgs()
    steps = [6 + 3 * i for i in range(len(cfgs))]
    srv = Server(0, 1 << 30, 4)
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
