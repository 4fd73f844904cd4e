"""Host-side mirror of the reference's step()/fluxes()/BCON() surface.

The reference advances global arrays with ``step()`` (3dmhdsub.h:31) and
refills ghosts with ``communicate()`` (3dmhdmpi.h:11).  :class:`Domain` holds
one device context and exposes the same operations on [nz][ny][nx] arrays:
``set_state``/``state`` replace direct access to the globals, ``step`` is
``step()``, ``fluxes`` is ``fluxes()``, ``bcon`` is ``BCON()``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .desc import Descriptor, RunConfig


class Domain:
    def __init__(self, cfg_or_desc):
        self.desc = cfg_or_desc if isinstance(cfg_or_desc, Descriptor) else Descriptor(cfg_or_desc)
        self.cfg = self.desc.cfg
        self.nf = 8 if self.cfg.lMag else 5
        self.shape = self.desc.shape
        h = C.c_void_p()
        rc = L.lib().mhd3d_create(C.cast(self.desc.ptr(), C.c_void_p), C.byref(h))
        if rc != L.OK:
            raise L.MhdError(rc, L.lib().mhd3d_last_error(None).decode())
        self.h = h

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            L.lib().mhd3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != L.OK:
            raise L.MhdError(rc, L.lib().mhd3d_last_error(self.h).decode())

    # -- state -------------------------------------------------------------
    def set_state(self, fields: dict):
        arrs = []
        for name in L.FIELDS[: self.nf]:
            a = np.ascontiguousarray(fields[name], dtype=np.float64)
            if a.shape != self.shape:
                raise ValueError(f"{name}: shape {a.shape} != {self.shape}")
            arrs.append(a)
        self._check(L.lib().mhd3d_set_state(self.h, L._fieldptrs(arrs)))

    def init_stratified(self, seed=12345, amp=1e-4, b0=1e-3, b_amp=1e-5):
        """Build the synthetic stratified state on the device
        (stratified_state_hashed is its host copy)."""
        from .initial import stratified_profiles
        T, rho = stratified_profiles(self.desc)
        T = np.ascontiguousarray(T)
        rho = np.ascontiguousarray(rho)
        dp = C.POINTER(C.c_double)
        self._check(L.lib().mhd3d_init_stratified(self.h, T.ctypes.data_as(dp), rho.ctypes.data_as(dp),
                                                  int(seed), float(amp), float(b0), float(b_amp)))

    def state(self) -> dict:
        arrs = [np.empty(self.shape) for _ in range(self.nf)]
        self._check(L.lib().mhd3d_get_state(self.h, L._fieldptrs(arrs)))
        return dict(zip(L.FIELDS[: self.nf], arrs))

    # -- the reference's operations ------------------------------------------
    def step(self, nsteps=1, dt=None):
        """step() x nsteps; CFL DT unless ``dt`` is given.  Returns the last DT."""
        d = C.c_double(0.0 if dt is None else float(dt))
        self._check(L.lib().mhd3d_step(self.h, int(nsteps), int(dt is not None), C.byref(d)))
        return d.value

    def stage(self, k, dt):
        self._check(L.lib().mhd3d_stage(self.h, int(k), float(dt)))

    def fluxes(self) -> dict:
        arrs = [np.empty(self.shape) for _ in range(self.nf)]
        self._check(L.lib().mhd3d_fluxes(self.h, L._fieldptrs(arrs)))
        return dict(zip(L.FLUXES[: self.nf], arrs))

    def bcon(self):
        self._check(L.lib().mhd3d_bcon(self.h))

    def cfl(self):
        info = L.CflInfo()
        self._check(L.lib().mhd3d_cfl(self.h, C.byref(info)))
        return info.dt, np.array(info.wmin[:]), info.umach

    def diagnostics(self):
        out = np.zeros(3)
        self._check(L.lib().mhd3d_diagnostics(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return {"mass": out[0], "energy": out[1], "max_divB": out[2]}

    def clock(self):
        n = C.c_int64()
        t = C.c_double()
        tc = C.c_double()
        self._check(L.lib().mhd3d_clock(self.h, C.byref(n), C.byref(t), C.byref(tc)))
        return n.value, t.value, tc.value

    def autotune(self, steps=2):
        """Pick the fastest streaming variant on the current state (state and
        clock unchanged).  Returns the trials as (waves, sched, planes, ms)."""
        rep = (C.c_double * (4 * 24))()
        n = C.c_int()
        self._check(L.lib().mhd3d_autotune(self.h, int(steps), rep, C.byref(n)))
        return [tuple(rep[4 * i: 4 * i + 4]) for i in range(n.value)]

    def variant(self):
        """(waves per SIMD, schedule bits, blocks per launch) in use."""
        w, s, b = C.c_int(), C.c_int(), C.c_int()
        self._check(L.lib().mhd3d_get_variant(self.h, C.byref(w), C.byref(s), C.byref(b)))
        return w.value, s.value, b.value

    def relax_state(self):
        """The RLAX statics (3dmhdsub.h:195-198): dict TFAC, TTM_MAX, TTM_MIN,
        ICALL and exited (the reference's exit point reached)."""
        out = (C.c_double * 3)()
        ic = C.c_int64()
        ex = C.c_int()
        self._check(L.lib().mhd3d_relax_state(self.h, out, C.byref(ic), C.byref(ex)))
        return {"TFAC": out[0], "TTM_MAX": out[1], "TTM_MIN": out[2], "ICALL": ic.value,
                "exited": ex.value}

    def timing(self, on=None):
        if on is not None:
            self._check(L.lib().mhd3d_timing_enable(self.h, int(bool(on))))
            return None
        t = L.Timing()
        self._check(L.lib().mhd3d_timing_read(self.h, C.byref(t)))
        return {"stage_ms_total": t.stage_ms_total, "stage_launches": t.stage_launches}

    def set_dt_guard(self, dt_min=L.REFERENCE_DT_MIN):
        """The reference driver's DT guard: steps after one with DT < dt_min
        do nothing and step() raises MhdError(EHALTED).  0 disables it."""
        self._check(L.lib().mhd3d_set_dt_guard(self.h, float(dt_min)))

    def synchronize(self):
        self._check(L.lib().mhd3d_synchronize(self.h))

    @property
    def device_bytes(self):
        return int(L.lib().mhd3d_device_bytes(self.h))
