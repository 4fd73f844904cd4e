"""The reference's dump/restart files (3dmhdio.h:20-72) through the library
(mhd3d_dump_data / mhd3d_load_data / mhd3d_dump_param / mhd3d_load_param).

`par1` records follow 3dmhd.cpp:271-289: par1[0..31] = ipar (3dmhdset.h:
27-48 and 3dmhd.cpp:371-392), par1[32..95] = par, then nIt at [14], DT at
[62], and at [63], [64] either TIMT, 0 (the last dump) or TIMI, TIMC.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from .desc import RunConfig

RECORD_FIELDS = ("RU", "RV", "RW", "TT", "RO", "BX", "BY", "BZ")   # 3dmhdio.h:24-33


def make_ipar(cfg: RunConfig, nCase=1, nCaseP=1, nTotal=0, nStep0=1, nStart=0, nPE=1, nPEy=1,
              nTube=0, nGrid=0, nCol=0, nRow=0):
    ipar = np.zeros(32)
    ipar[0], ipar[1], ipar[2], ipar[3], ipar[4] = nCase, nCaseP, nTotal, nStep0, nStart
    ipar[5], ipar[6], ipar[7] = cfg.ixCon, cfg.iyCon, cfg.izCon
    ipar[8], ipar[9] = cfg.iTCon, cfg.iBCon
    ipar[10], ipar[11], ipar[12] = cfg.nPx, cfg.nPy_global or cfg.nPy, cfg.nPz_global or cfg.nPz
    ipar[13] = nPE
    ipar[15], ipar[16], ipar[17], ipar[18] = nTube, nGrid, nCol, nRow
    ipar[19] = cfg.iD
    ipar[20] = nPEy
    return ipar.astype(np.float32).astype(np.float64)   # stored through float (3dmhd.cpp:271)


def make_par1(cfg: RunConfig, nIt, DT, TIMI=0.0, TIMC=0.0, TIMT=None, last=False, **ipar_kw):
    par1 = np.zeros(96)
    par1[:32] = make_ipar(cfg, **ipar_kw)
    par1[32:96] = np.asarray(cfg.par, dtype=np.float64)[:64]
    par1[14] = nIt
    par1[62] = DT
    if last:
        par1[63] = TIMI + TIMC if TIMT is None else TIMT
        par1[64] = 0.0
    else:
        par1[63] = TIMI
        par1[64] = TIMC
    return par1


def dump_param(path, idx, par1):
    a = np.ascontiguousarray(par1, dtype=np.float64)
    assert a.shape == (96,)
    rc = L.lib().mhd3d_dump_param(os.fsencode(path), int(idx), a.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != L.OK:
        raise L.MhdError(rc, f"cannot write record {idx} of {path}")


def load_param(path, idx):
    a = np.zeros(96)
    rc = L.lib().mhd3d_load_param(os.fsencode(path), int(idx), a.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != L.OK:
        raise L.MhdError(rc, f"cannot read record {idx} of {path}")
    return a


def dump_data(domain, path, idx):
    domain._check(L.lib().mhd3d_dump_data(domain.h, os.fsencode(path), int(idx)))


def load_data(domain, path, idx):
    domain._check(L.lib().mhd3d_load_data(domain.h, os.fsencode(path), int(idx)))
