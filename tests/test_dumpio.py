""".par records of the reference's dump format (3dmhdio.h:37-55), written
and read by the library's host code (no GPU).  The .dat0 state records need
a device context: tests/test_gpu_dumpio.py."""
import os

import numpy as np
import pytest

from mhd3d import _lib as L
from mhd3d.dumpio import dump_param, load_param, make_par1
from mhd3d.initial import stratified_config

pytestmark = pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="libmhd3d.so not built")


def test_param_records_roundtrip_and_layout(tmp_path):
    cfg = stratified_config(48, 40, 20, lMag=True)
    path = str(tmp_path / "run.par")
    p0 = make_par1(cfg, nIt=0, DT=0.0)
    p1 = make_par1(cfg, nIt=10, DT=1.25e-4, TIMI=0.5, TIMC=0.00125)
    p2 = make_par1(cfg, nIt=20, DT=1.5e-4, TIMI=0.5, TIMC=0.0027, last=True)
    dump_param(path, 0, p0)
    dump_param(path, 2, p2)   # out of order: record 1 stays a hole until written
    dump_param(path, 1, p1)
    raw = np.fromfile(path, dtype="<f8")
    assert raw.size == 3 * 96                       # 96 doubles per record, nothing else
    assert np.array_equal(raw.reshape(3, 96)[1], p1)
    for k, p in enumerate((p0, p1, p2)):
        assert np.array_equal(load_param(path, k), p)
    # the layout of 3dmhd.cpp:271-289
    assert p1[10] == 48 and p1[11] == 40 and p1[12] == 20      # nPx, nPy, nPz
    assert p1[14] == 10 and p1[62] == 1.25e-4                   # nIt, DT
    assert p1[63] == 0.5 and p1[64] == 0.00125                  # TIMI, TIMC
    assert p2[63] == 0.5 + 0.0027 and p2[64] == 0.0             # TIMT, 0 on the last dump
    par = np.asarray(cfg.par, dtype=np.float64)[:64]
    keep = [k for k in range(32, 96) if k not in (62, 63, 64)]   # DT, times overwrite par[30..32]
    assert np.array_equal(p1[keep], par[[k - 32 for k in keep]])


def test_missing_param_record_is_an_error(tmp_path):
    path = str(tmp_path / "short.par")
    dump_param(path, 0, np.zeros(96))
    with pytest.raises(L.MhdError):
        load_param(path, 3)
