"""ctypes binding of the CPU restatement (oracle/build/libmhd_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libmhd_oracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def build_variant(out_path, defines=()):
    """Compile the restatement with extra -D flags (planted-mistake builds)."""
    cmd = ["gcc", "-O2", "-fPIC", "-std=c11", "-fopenmp", "-ffp-contract=off", "-shared",
           *("-D" + d for d in defines), os.path.join(HERE, "mhd_oracle.c"), "-o", out_path, "-lm"]
    subprocess.run(cmd, check=True)
    return out_path


def load(path):
    """Open a build of the restatement and declare its entry points."""
    L = C.CDLL(path)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.orc_create.restype = vp
    L.orc_create.argtypes = [vp]
    L.orc_destroy.argtypes = [vp]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set.argtypes = [vp, C.c_char_p, dp]
    L.orc_get.argtypes = [vp, C.c_char_p, dp]
    for fn in ("orc_communicate", "orc_fluxes", "orc_bcon"):
        getattr(L, fn).argtypes = [vp]
    L.orc_cfl.restype = C.c_double
    L.orc_cfl.argtypes = [vp, dp, dp]
    L.orc_stage.argtypes = [vp, C.c_int, C.c_double]
    L.orc_step.restype = C.c_double
    L.orc_step.argtypes = [vp, C.c_int, C.c_double]
    L.orc_clock.argtypes = [vp, C.POINTER(C.c_longlong), dp, dp]
    L.orc_set_time.argtypes = [vp, C.c_double, C.c_double]
    L.orc_relax_state.restype = C.c_int
    L.orc_relax_state.argtypes = [vp, dp, C.POINTER(C.c_longlong), C.POINTER(C.c_int), dp, dp, dp]
    L.orc_ran2.restype = C.c_double
    L.orc_ran2.argtypes = [C.POINTER(C.c_int)]
    L.orc_static_profile.restype = C.c_int
    L.orc_static_profile.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp, dp]
    L.orc_static_rnd.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, dp]
    L.orc_rempel_profile.restype = C.c_int
    L.orc_rempel_profile.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp, dp]
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


FIELDS = ("RU", "RV", "RW", "RO", "TT", "BX", "BY", "BZ")
FLUXES = ("FU", "FV", "FW", "FR", "FT", "WW1", "WW2", "WW3")


class Oracle:
    """One domain of the CPU restatement; mirrors the reference's globals."""

    def __init__(self, desc, threads: int | None = None):
        self.desc = desc  # keeps the metric arrays alive
        self.shape = desc.shape
        if threads:
            lib().orc_set_threads(int(threads))
        self.h = lib().orc_create(C.cast(desc.ptr(), C.c_void_p))
        if not self.h:
            raise RuntimeError("orc_create failed")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        if lib().orc_set(self.h, name.encode(), _dp(a)) != 0:
            raise KeyError(name)

    def get(self, name):
        a = np.empty(self.shape)
        if lib().orc_get(self.h, name.encode(), _dp(a)) != 0:
            raise KeyError(name)
        return a

    def set_state(self, fields: dict):
        for k, v in fields.items():
            self.set(k, v)

    def state(self, mag=False):
        names = FIELDS if mag else FIELDS[:5]
        return {k: self.get(k) for k in names}

    def communicate(self):
        lib().orc_communicate(self.h)

    def fluxes(self):
        lib().orc_fluxes(self.h)

    def bcon(self):
        lib().orc_bcon(self.h)

    def cfl(self):
        w = np.zeros(4)
        um = C.c_double()
        dt = lib().orc_cfl(self.h, _dp(w), C.byref(um))
        return dt, w, um.value

    def stage(self, k, dt):
        lib().orc_stage(self.h, int(k), float(dt))

    def step(self, fixed_dt=False, dt=0.0):
        return lib().orc_step(self.h, int(bool(fixed_dt)), float(dt))

    def clock(self):
        n = C.c_longlong()
        t = C.c_double()
        tc = C.c_double()
        lib().orc_clock(self.h, C.byref(n), C.byref(t), C.byref(tc))
        return n.value, t.value, tc.value

    def set_time(self, dt, timt):
        """The globals DT and TIMT that fluxes() reads."""
        lib().orc_set_time(self.h, float(dt), float(timt))

    def relax_state(self):
        """RLAX statics and last profiles: dict TFAC, TTM_MAX, TTM_MIN, ICALL,
        exited, CORRECT, FCONM, FRADM."""
        out = np.zeros(3)
        ic = C.c_longlong()
        ex = C.c_int()
        nz = self.shape[0]
        prof = [np.zeros(nz) for _ in range(3)]
        lib().orc_relax_state(self.h, _dp(out), C.byref(ic), C.byref(ex), *(_dp(p) for p in prof))
        return {"TFAC": out[0], "TTM_MAX": out[1], "TTM_MIN": out[2], "ICALL": ic.value,
                "exited": ex.value, "CORRECT": prof[0], "FCONM": prof[1], "FRADM": prof[2]}


def static_profile(nPz, THETA, GRAV, zMax):
    T = np.zeros(nPz)
    R = np.zeros(nPz)
    if lib().orc_static_profile(nPz, THETA, GRAV, zMax, _dp(T), _dp(R)) != 0:
        raise RuntimeError("bsstep failed")
    return T, R


def rempel_profile(nPz, THETA, REPR, zMax):
    """lRem background of Static(): T, R, RK, DRK and (TU, DZTU, DZTB)."""
    T, R, RK, DRK = (np.zeros(nPz) for _ in range(4))
    walls = np.zeros(3)
    if lib().orc_rempel_profile(nPz, THETA, REPR, zMax, _dp(T), _dp(R), _dp(RK), _dp(DRK), _dp(walls)) != 0:
        raise RuntimeError("ODEINT failed")
    return T, R, RK, DRK, walls


def static_rnd(nPx, nPy, nzg, nPEy, nPEz, AMPT):
    out = np.zeros((nzg + 2, nPy + 2, nPx + 2))
    lib().orc_static_rnd(nPx, nPy, nzg, nPEy, nPEz, AMPT, _dp(out))
    return out


def ran2_sequence(n, idum=-62659):
    st = (C.c_int * 99)()
    st[0] = idum
    return np.array([lib().orc_ran2(st) for _ in range(n)])
