"""The single-domain autotuner (mhd3d_autotune) on the GPU: it times every
streaming variant on the current state and keeps the fastest, leaving the
state, the clock and the results untouched -- after it, steps are
bit-identical to those of a domain that never ran it (FP64 and mixed; the
mixed form against itself, as its variants compute each point the same)."""
import numpy as np
import pytest

from mhd3d import _lib as L
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.initial import stratified_config, stratified_state

pytestmark = pytest.mark.gpu


def _same(a, b):
    return all(np.array_equal(a[k].view(np.int64), b[k].view(np.int64)) for k in a)


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_changes_nothing_but_the_variant(precision):
    print(f"autotune precision={precision}", flush=True)
    desc = Descriptor(stratified_config(64, 48, 40, lMag=True, lRot=True, precision=precision))
    st = stratified_state(desc, seed=31)
    dom = Domain(desc)
    dom.set_state(st)
    dom.step(2)
    before, clock = dom.state(), dom.clock()
    trials = dom.autotune(steps=2)
    assert len(trials) == 24
    assert all(t[3] > 0 for t in trials), trials
    w, s, b = dom.variant()
    assert (w, s) in {(t[0], t[1]) for t in trials} and b >= 1
    assert _same(dom.state(), before) and dom.clock() == clock
    dt = dom.step(5)
    ref = Domain(desc)
    ref.set_state(st)
    ref.step(2)
    assert ref.step(5) == dt
    assert _same(dom.state(), ref.state())
    assert dom.clock() == ref.clock()
    dom.close()
    ref.close()


def test_autotune_refuses_branches():
    import test_gpu_ext as tx
    dom = Domain(Descriptor(tx._cfg((16, 12, 8), "lShr")))
    with pytest.raises(L.MhdError) as ei:
        dom.autotune()
    assert ei.value.code == L.EUNSUPPORTED
    dom.close()
