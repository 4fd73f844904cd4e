"""State records of the reference's .dat0 dump (3dmhdio.h:20-72): the bytes
on disk follow save() (3dmhdutils.h:105-110) field by field, a dump read back
restores the state, and a run restarted from a dump continues bit-for-bit."""
import numpy as np
import pytest

from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.dumpio import RECORD_FIELDS, dump_data, load_data
from mhd3d.initial import stratified_config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lMag", [False, True])
def test_dat0_layout_and_restart(tmp_path, lMag):
    d = Descriptor(stratified_config(40, 24, 12, lMag=lMag, lRot=True))
    nz, ny, nx = d.shape
    nf = 8 if lMag else 5
    core = (nx - 2) * (ny - 2) * (nz - 2)
    dom = Domain(d)
    dom.init_stratified(seed=11)
    dom.step(3)
    s3 = dom.state()
    path = str(tmp_path / "run.dat0.000")
    dump_data(dom, path, 0)
    dom.step(2)
    dump_data(dom, path, 1)
    s5 = dom.state()
    raw = np.fromfile(path, dtype="<f8")
    assert raw.size == 2 * nf * core
    recs = raw.reshape(2, nf, nz - 2, ny - 2, nx - 2)
    for r, name in enumerate(RECORD_FIELDS[:nf]):
        assert np.array_equal(recs[0, r], s3[name][1:-1, 1:-1, 1:-1]), name
        assert np.array_equal(recs[1, r], s5[name][1:-1, 1:-1, 1:-1]), name
    # restart from dump 0 in a fresh context holding the same initial ghosts
    dom2 = Domain(d)
    dom2.init_stratified(seed=11)
    load_data(dom2, path, 0)
    back = dom2.state()
    for name in s3:
        assert np.array_equal(back[name], s3[name]), name
    dom2.step(2)
    for name, v in dom2.state().items():
        assert np.array_equal(v.view(np.int64), s5[name].view(np.int64)), name
