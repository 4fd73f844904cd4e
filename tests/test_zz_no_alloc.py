"""mhd3d_step_noalloc through the C ABI: the same results as mhd3d_step, bit
for bit, on two calls in a row (the second on a smaller domain, so the
shared partial records hold the first call's contents), with CFL and fixed
DT; grids that need more partial records than the array holds are refused
with a message."""
import ctypes as C

import numpy as np
import pytest

from mhd3d import _lib as L
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.initial import stratified_config, stratified_state

pytestmark = pytest.mark.gpu


def _noalloc(dom, nsteps, dt=None):
    d = C.c_double(0.0 if dt is None else float(dt))
    rc = L.lib().mhd3d_step_noalloc(dom.h, int(nsteps), int(dt is not None), C.byref(d))
    return rc, d.value


def _same(a, b):
    return all(np.array_equal(a[k].view(np.int64), b[k].view(np.int64)) for k in a)


@pytest.mark.parametrize("fixed", [False, True])
def test_noalloc_step_matches_step_on_two_calls(fixed):
    cases = [(stratified_config(64, 48, 24, lMag=True, lRot=True), 5),
             (stratified_config(32, 24, 12), 3)]          # second call: a smaller domain
    for cfg, n in cases:
        desc = Descriptor(cfg)
        st = stratified_state(desc, seed=31)
        ref = Domain(desc)
        ref.set_state(st)
        dt = ref.cfl()[0] if fixed else None
        dt_ref = ref.step(n, dt=dt)
        dom = Domain(desc)
        dom.set_state(st)
        rc, dt_got = _noalloc(dom, n, dt)
        assert rc == L.OK, L.lib().mhd3d_last_error(dom.h).decode()
        assert dt_got == dt_ref
        assert dom.clock() == ref.clock()
        assert _same(dom.state(), ref.state()), (cfg.nPx, cfg.nPy, cfg.nPz)
        # the context's own step continues from the same state and clock
        assert dom.step(2, dt=dt) == ref.step(2, dt=dt)
        assert _same(dom.state(), ref.state())
        dom.close()
        ref.close()


def test_noalloc_refuses_grids_beyond_its_records():
    # 2048 x 1024 x 5 hydro needs 32 * 256 * 5 = 40960 standalone TSTEP records
    dom = Domain(Descriptor(stratified_config(2048, 1024, 5)))
    rc, _ = _noalloc(dom, 1)
    assert rc == L.EINVAL
    assert "32768" in L.lib().mhd3d_last_error(dom.h).decode()
    assert dom.clock()[0] == 0
    dom.close()
