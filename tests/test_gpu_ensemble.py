"""Ensembles: members of different sizes, flags and parameters advanced in one
persistent launch per stage.  Bar: every member bit-identical to the same
domain advanced alone (FP64 and mixed forms), with the same DT and clock;
graph replays, kernel variants and the autotuner change nothing."""
import numpy as np
import pytest

from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.ensemble import Ensemble
from mhd3d.initial import stratified_config, stratified_state

pytestmark = pytest.mark.gpu

MEMBERS = [
    dict(n=(24, 16, 9), lMag=False, lRot=False, Ra=3e4),
    dict(n=(40, 24, 12), lMag=True, lRot=True, Ra=1e5),
    dict(n=(33, 17, 7), lMag=False, lRot=True, Ra=2e5),
    dict(n=(16, 32, 10), lMag=True, lRot=False, Ra=5e4),
    dict(n=(64, 40, 20), lMag=True, lRot=True, Ra=1e6, Pm=0.5),
]


def _members(precision):
    descs, states = [], []
    for k, m in enumerate(MEMBERS):
        kw = {key: v for key, v in m.items() if key != "n"}
        cfg = stratified_config(*m["n"], precision=precision, **kw)
        d = Descriptor(cfg)
        descs.append(d)
        states.append(stratified_state(d, seed=100 + k))
    return descs, states


def _alone(desc, st, nsteps):
    dom = Domain(desc)
    dom.set_state(st)
    dt = dom.step(nsteps)
    return dom.state(), dt, dom.clock()


def _same(a, b):
    return all(np.array_equal(a[k].view(np.int64), b[k].view(np.int64)) for k in a)


@pytest.mark.parametrize("precision", [0, 1])
def test_members_bit_identical_to_alone(precision):
    descs, states = _members(precision)
    ens = Ensemble(descs)
    for i, st in enumerate(states):
        ens.set_state(i, st)
    ens.step(4)
    ens.step(3)
    for i, (d, st) in enumerate(zip(descs, states)):
        ref, dt, (n, timt, timc) = _alone(d, st, 7)
        got = ens.state(i)
        assert _same(got, ref), (precision, i)
        n_e, timt_e, timc_e, dt_e = ens.clock(i)
        assert n_e == n == 7 and timc_e == timc and dt_e == dt, (i, n_e, timc_e, timc, dt_e, dt)


def test_graph_replays_match_eager():
    descs, states = _members(0)
    eager = Ensemble(descs)
    graph = Ensemble(descs)
    graph.graph(True)
    for i, st in enumerate(states):
        eager.set_state(i, st)
        graph.set_state(i, st)
    for rep in range(4):
        eager.step(2)
        graph.step(2)
        # eager work of the program between the replays
        for i in range(len(descs)):
            assert _same(graph.state(i), eager.state(i)), (rep, i)


@pytest.mark.parametrize("variant", [(1, 0, 0), (1, 3, 8), (2, 2, 16), (2, 1, 32)])
def test_variants_do_not_change_results(variant):
    descs, states = _members(1)
    ref = Ensemble(descs)
    ens = Ensemble(descs)
    ens.variant(*variant)
    for i, st in enumerate(states):
        ref.set_state(i, st)
        ens.set_state(i, st)
    ref.step(3)
    ens.step(3)
    for i in range(len(descs)):
        assert _same(ens.state(i), ref.state(i)), (variant, i)


def test_autotune_restores_state_and_picks_a_variant():
    descs, states = _members(0)
    ens = Ensemble(descs)
    for i, st in enumerate(states):
        ens.set_state(i, st)
    ens.step(2)
    before = [ens.state(i) for i in range(len(descs))]
    clocks = [ens.clock(i) for i in range(len(descs))]
    best, trials = ens.autotune(steps=2)
    assert len(trials) == 32 and all(t["ms_per_step"] > 0 for t in trials)
    assert best in [(t["waves"], t["sched"], t["chunk"]) for t in trials]
    for i in range(len(descs)):
        assert _same(ens.state(i), before[i]), i
        assert ens.clock(i) == clocks[i]
    ens.step(5)
    for i, (d, st) in enumerate(zip(descs, states)):
        ref, _, _ = _alone(d, st, 7)
        assert _same(ens.state(i), ref), i


def test_set_state_resets_one_member():
    descs, states = _members(0)
    ens = Ensemble(descs)
    for i, st in enumerate(states):
        ens.set_state(i, st)
    ens.step(3)
    ens.set_state(2, states[2])
    ens.step(2)
    ref, dt, (n, _, timc) = _alone(descs[2], states[2], 2)
    assert _same(ens.state(2), ref)
    assert ens.clock(2)[0] == 2 and ens.clock(2)[2] == timc
    ref1, _, _ = _alone(descs[1], states[1], 5)
    assert _same(ens.state(1), ref1)
