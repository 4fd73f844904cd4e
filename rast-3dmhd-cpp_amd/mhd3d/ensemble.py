"""Ensembles of independent domains advanced together (mhd3d_ensemble_* of
include/mhd3d.h): one persistent launch per RK3 stage covers every member,
whatever its size, flags and parameters; each member keeps its own CFL time
step and clock, and advances bit-for-bit as it would alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .desc import Descriptor


class Ensemble:
    def __init__(self, members):
        self.descs = [m if isinstance(m, Descriptor) else Descriptor(m) for m in members]
        n = len(self.descs)
        arr = (C.c_void_p * n)(*[C.cast(d.ptr(), C.c_void_p) for d in self.descs])
        h = C.c_void_p()
        rc = L.lib().mhd3d_ensemble_create(arr, n, C.byref(h))
        if rc != L.OK:
            raise L.MhdError(rc, L.lib().mhd3d_ensemble_last_error(None).decode())
        self.h = h
        self.n = n

    def close(self):
        if getattr(self, "h", None):
            L.lib().mhd3d_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != L.OK:
            raise L.MhdError(rc, L.lib().mhd3d_ensemble_last_error(self.h).decode())

    def _nf(self, i):
        return 8 if self.descs[i].cfg.lMag else 5

    @property
    def cells(self):
        return int(L.lib().mhd3d_ensemble_cells(self.h))

    def set_state(self, i, fields: dict):
        shape = self.descs[i].shape
        arrs = [np.ascontiguousarray(fields[k], dtype=np.float64) for k in L.FIELDS[: self._nf(i)]]
        for k, a in zip(L.FIELDS, arrs):
            if a.shape != shape:
                raise ValueError(f"member {i} {k}: shape {a.shape} != {shape}")
        self._check(L.lib().mhd3d_ensemble_set_state(self.h, int(i), L._fieldptrs(arrs)))

    def state(self, i) -> dict:
        shape = self.descs[i].shape
        arrs = [np.empty(shape) for _ in range(self._nf(i))]
        self._check(L.lib().mhd3d_ensemble_get_state(self.h, int(i), L._fieldptrs(arrs)))
        return dict(zip(L.FIELDS[: self._nf(i)], arrs))

    def step(self, nsteps=1):
        self._check(L.lib().mhd3d_ensemble_step(self.h, int(nsteps)))

    def clock(self, i):
        n = C.c_int64()
        t = C.c_double()
        tc = C.c_double()
        dt = C.c_double()
        self._check(L.lib().mhd3d_ensemble_clock(self.h, int(i), C.byref(n), C.byref(t),
                                                 C.byref(tc), C.byref(dt)))
        return n.value, t.value, tc.value, dt.value

    def graph(self, on=True):
        self._check(L.lib().mhd3d_ensemble_graph(self.h, int(bool(on))))

    def variant(self, waves=None, sched=None, chunk=None):
        w, s, c = C.c_int(), C.c_int(), C.c_int()
        self._check(L.lib().mhd3d_ensemble_get_variant(self.h, C.byref(w), C.byref(s), C.byref(c)))
        if waves is None and sched is None and chunk is None:
            return w.value, s.value, c.value
        self._check(L.lib().mhd3d_ensemble_set_variant(
            self.h, int(w.value if waves is None else waves), int(s.value if sched is None else sched),
            int(c.value if chunk is None else chunk)))
        return self.variant()

    def autotune(self, steps=3):
        rep = np.zeros(4 * 32)
        n = C.c_int()
        self._check(L.lib().mhd3d_ensemble_autotune(self.h, int(steps),
                                                    rep.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)))
        trials = [dict(waves=int(rep[4 * k]), sched=int(rep[4 * k + 1]), chunk=int(rep[4 * k + 2]),
                       ms_per_step=float(rep[4 * k + 3])) for k in range(n.value)]
        return self.variant(), trials

    def set_dt_guard(self, dt_min=L.REFERENCE_DT_MIN, member=None):
        """The DT guard for every member, or for one member."""
        if member is None:
            self._check(L.lib().mhd3d_ensemble_set_dt_guard(self.h, float(dt_min)))
        else:
            self._check(L.lib().mhd3d_ensemble_set_dt_guard_member(self.h, int(member), float(dt_min)))

    def halted(self, i):
        h = C.c_int()
        self._check(L.lib().mhd3d_ensemble_halted(self.h, int(i), C.byref(h)))
        return bool(h.value)

    def synchronize(self):
        self._check(L.lib().mhd3d_ensemble_synchronize(self.h))
