"""Descriptor checks of the off-default branches (CPU: mhd3d_create validates
the descriptor before its first HIP call, so a refused descriptor returns
without a device).

Refused: iD outside 0..2, missing coordinate arrays, RLAX without the z
rank layout of the reference run its scan reproduces (relax_nPEz >= 2
dividing nPz, at least 3 planes per rank, 3dmhdsub.h:354-421), and in
ensembles of either form the heat loss and the plume with TC != 0, whose
factors of TIMT the host forms before every step (3dmhdsub.h:446-460,752-757;
ensembles check their members before any device call).
"""
import ctypes as C
import dataclasses
import os

import pytest

from mhd3d import _lib as L
from mhd3d.desc import Descriptor
from mhd3d.initial import stratified_config


def _create(cfg, null_zee=False):
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libmhd3d.so not built")
    desc = Descriptor(cfg)
    if null_zee:
        desc.struct.ZEE = C.POINTER(C.c_double)()
    h = C.c_void_p()
    rc = L.lib().mhd3d_create(C.cast(desc.ptr(), C.c_void_p), C.byref(h))
    msg = L.lib().mhd3d_last_error(None)
    assert not h.value
    return rc, (msg or b"").decode()


def _base(**kw):
    return dataclasses.replace(stratified_config(12, 10, 8), **kw)


def test_rlax_needs_the_rank_layout():
    cfg = _base()
    par = cfg.par.copy()
    par[59] = 0.1
    for P in (0, 1, 3, 4):                  # nPz = 8: 3 does not divide it, 4 leaves 2 planes a rank
        rc, msg = _create(dataclasses.replace(cfg, par=par, relax_nPEz=P))
        assert rc == L.EINVAL and "relax_nPEz" in msg, P


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_refuses_time_dependent_branches(precision):
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libmhd3d.so not built")
    cfg = _base(precision=precision, iTCon=2, TIMI=0.001)
    par = cfg.par.copy()
    par[11], par[12], par[13], par[14], par[15] = 0.5, 0.002, 1.2, 0.9, 0.6   # heat loss, TC != 0
    par[52], par[54] = 0.01, 0.8
    desc = Descriptor(dataclasses.replace(cfg, par=par))
    arr = (C.c_void_p * 1)(C.cast(desc.ptr(), C.c_void_p))
    h = C.c_void_p()
    rc = L.lib().mhd3d_ensemble_create(arr, 1, C.byref(h))
    msg = (L.lib().mhd3d_ensemble_last_error(None) or b"").decode()
    assert rc == L.EUNSUPPORTED and "TIMT" in msg and not h.value


def test_bad_iD_and_missing_coordinates():
    rc, _ = _create(_base(iD=3))
    assert rc == L.EINVAL
    cfg = _base(iD=2)
    par = cfg.par.copy()
    par[56] = 0.5
    rc, msg = _create(dataclasses.replace(cfg, par=par), null_zee=True)
    assert rc == L.EINVAL and "ZEE" in msg
