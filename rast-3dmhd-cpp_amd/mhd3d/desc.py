"""ctypes mirror of ``mhd3d_desc`` (include/mhd3d.h) and the run-descriptor builder.

The reference hard-codes its run in compile-time macros (3dmhdparam.h:8-123)
and ``setup()`` (3dmhdset.h:7-24), then derives every scalar in ``set()``
(3dmhd.cpp:353-495) and the grid metrics in ``xJacobi/yJacobi/zJacobi``
(3dmhdphsic.h:79-156).  :func:`build_desc` restates ``set()`` for one domain,
keeping the float-literal roundings of the C++ source (``1.0f/3.0f`` etc.)
by doing those operations in float32, so the descriptor carries the same
doubles the reference's globals hold.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


def f32(x) -> float:
    """Value of a C++ float literal/expression, as the double it promotes to."""
    return float(F32(x))


def fdiv(a, b) -> float:
    """``a##f / b##f`` evaluated in float, promoted to double."""
    return float(F32(a) / F32(b))


class MhdDesc(C.Structure):
    """Field-for-field mirror of ``struct mhd3d_desc``."""

    _fields_ = [
        ("nPx", C.c_int32), ("nPy", C.c_int32), ("nPz", C.c_int32),
        ("flag_lRot", C.c_int32), ("flag_lMag", C.c_int32), ("flag_lPot", C.c_int32),
        ("flag_lRem", C.c_int32), ("flag_lShr", C.c_int32),
        ("flag_ixCon", C.c_int32), ("flag_iyCon", C.c_int32), ("flag_izCon", C.c_int32),
        ("flag_iTCon", C.c_int32), ("flag_iBCon", C.c_int32), ("flag_iD", C.c_int32),
        ("precision", C.c_int32), ("relax_nPEz", C.c_int32),
        ("C13", C.c_double), ("C23", C.c_double), ("C43", C.c_double),
        ("GAM1", C.c_double), ("GAM2", C.c_double), ("GAM3", C.c_double),
        ("ZETA1", C.c_double), ("ZETA2", C.c_double),
        ("SFF", C.c_double),
        ("hx", C.c_double), ("h2x", C.c_double), ("hy", C.c_double),
        ("h2y", C.c_double), ("hz", C.c_double), ("h2z", C.c_double),
        ("ORZF", C.c_double),
        ("RE", C.c_double), ("ORE", C.c_double), ("REPR", C.c_double),
        ("CV", C.c_double), ("OCV", C.c_double), ("GAMMA", C.c_double),
        ("THETA", C.c_double), ("GRAV", C.c_double),
        ("OMX", C.c_double), ("OMZ", C.c_double),
        ("RM", C.c_double), ("ORM", C.c_double), ("OBETA", C.c_double),
        ("TP", C.c_double), ("TB", C.c_double), ("TC", C.c_double),
        ("RLAX", C.c_double), ("TSTART", C.c_double), ("TOFF", C.c_double),
        ("TIMI", C.c_double),
        ("dxxdx", C.POINTER(C.c_double)), ("d2xxdx2", C.POINTER(C.c_double)),
        ("DDX", C.POINTER(C.c_double)),
        ("dyydy", C.POINTER(C.c_double)), ("d2yydy2", C.POINTER(C.c_double)),
        ("DDY", C.POINTER(C.c_double)),
        ("dzzdz", C.POINTER(C.c_double)), ("d2zzdz2", C.POINTER(C.c_double)),
        ("DDZ", C.POINTER(C.c_double)),
        ("RKAPA", C.POINTER(C.c_double)), ("DKAPA", C.POINTER(C.c_double)),
        ("DH", C.c_double), ("DV", C.c_double), ("zMax", C.c_double),
        ("TU", C.c_double), ("DZTU", C.c_double), ("DZTB", C.c_double),
        ("XP", C.c_double), ("YP", C.c_double), ("ZP", C.c_double),
        ("QFH", C.c_double), ("HH", C.c_double), ("ORZD", C.c_double),
        ("EXX", C.POINTER(C.c_double)), ("WYY", C.POINTER(C.c_double)),
        ("ZEE", C.POINTER(C.c_double)),
    ]


def default_par() -> np.ndarray:
    """``par[64]`` as ``setup()`` leaves it (3dmhdset.h:7-24)."""
    par = np.zeros(64)
    par[0] = 75
    par[1] = f32(0.2)
    par[2] = 5
    par[3] = 10
    par[11] = f32(0.1)
    par[16] = f32(6.0)
    par[17] = f32(6.0)
    par[18] = f32(2.0)
    par[19] = f32(1.0e-03)
    par[53] = fdiv(5.0, 3.0)
    return par


@dataclass
class Metrics:
    """1-D grid arrays of one domain (3dmhddata.h:20-22, 32)."""

    EXX: np.ndarray
    dxxdx: np.ndarray
    d2xxdx2: np.ndarray
    DDX: np.ndarray
    WYY: np.ndarray
    dyydy: np.ndarray
    d2yydy2: np.ndarray
    DDY: np.ndarray
    ZEE: np.ndarray
    dzzdz: np.ndarray
    d2zzdz2: np.ndarray
    DDZ: np.ndarray
    RKAPA: np.ndarray
    DKAPA: np.ndarray


def _uniform_axis(n_interior: int, n_global: int, smax: float, ghosts_zero: bool):
    """Uniform grid (nGrid=0, mkGrid 3dmhdphsic.h:37-41) on one axis.

    ``xJacobi`` (3dmhdphsic.h:79-91) fills interior points only and leaves the
    ghost entries at zero; ``yJacobi/zJacobi`` (:93-156) fill every entry.
    Position index i (0..n_interior-1) maps to code coordinate float(i)*OR
    with OR = 1.0f/float(n_global-1)."""
    n = n_interior + 2
    OR = fdiv(1.0, float(n_global - 1))
    s = np.zeros(n)
    du = np.zeros(n)
    d2 = np.zeros(n)
    dd = np.zeros(n)
    lo, hi = (1, n - 1) if ghosts_zero else (0, n)
    for i in range(lo, hi):
        k = min(max(i - 1, 0), n_interior - 1)
        code = float(k) * OR
        s[i] = code * smax
        du[i] = 1.0 / smax
        d2[i] = 0.0
        dd[i] = OR * (1.0 / du[i])
    return s, du, d2, dd


def tanh_axis(n_interior: int, smax: float, beta: float):
    """Builder's z grid, symmetric tanh stretching clustered at both walls:
    z(s) = smax/2 * (1 + tanh(beta*(2s-1))/tanh(beta)), s = k/(n-1).
    Returns the metric arrays in the reference's convention: dzzdz = ds/dz,
    d2zzdz2 = d2s/dz2, DDZ = OR * dz/ds (3dmhdphsic.h:155)."""
    n = n_interior + 2
    OR = fdiv(1.0, float(n_interior - 1))
    A = smax * beta / math.tanh(beta)
    z = np.zeros(n)
    du = np.zeros(n)
    d2 = np.zeros(n)
    dd = np.zeros(n)
    for i in range(n):
        k = min(max(i - 1, 0), n_interior - 1)
        s = float(k) * OR
        u = beta * (2.0 * s - 1.0)
        sech2 = 1.0 / math.cosh(u) ** 2
        dzds = A * sech2
        d2zds2 = -4.0 * A * beta * sech2 * math.tanh(u)
        z[i] = 0.5 * smax * (1.0 + math.tanh(u) / math.tanh(beta))
        du[i] = 1.0 / dzds
        d2[i] = -d2zds2 / dzds ** 3
        dd[i] = OR * (1.0 / du[i])
    return z, du, d2, dd


@dataclass
class RunConfig:
    """One run: sizes, the compile-time flags of 3dmhdparam.h and ``par[64]``."""

    nPx: int
    nPy: int
    nPz: int                      # z planes in this domain incl. both walls
    nPz_global: int | None = None  # nPz of 3dmhdparam.h (hz, ORZ); default nPz
    nPy_global: int | None = None
    lRot: bool = False
    lMag: bool = False
    lPot: bool = False
    lRem: bool = False
    lShr: bool = False
    ixCon: int = 0
    iyCon: int = 0
    izCon: int = 1
    iTCon: int = 4
    iBCon: int = 0
    iD: int = 0
    par: np.ndarray = field(default_factory=default_par)
    TB: float = 0.0
    # ("uniform",), ("tanh", beta) or ("arrays", ZEE, dzzdz, d2zzdz2, DDZ): the
    # z metric arrays of a caller that maps its own grid (zJacobi's outputs,
    # 3dmhdphsic.h:127-156), each of length nPz + 2
    zgrid: tuple = ("uniform",)
    precision: int = 0
    TIMI: float = 0.0
    # lRem: the wall temperatures Static() leaves in TU, DZTU, DZTB
    # (3dmhdphsic.h:590-592) and, since set() leaves RKAPA/DKAPA to Static()
    # under lRem (3dmhd.cpp:477), the conductivity profile (RKAPA, DKAPA)
    TU: float = 0.0
    DZTU: float = 0.0
    DZTB: float = 0.0
    kappa: tuple | None = None
    # RLAX > 0: the z ranks of the reference run whose relaxation scan
    # (3dmhdsub.h:354-421) is reproduced
    relax_nPEz: int = 0


class Descriptor:
    """A filled ``MhdDesc`` plus the numpy arrays it points at (kept alive)."""

    def __init__(self, cfg: RunConfig):
        self.cfg = cfg
        self.metrics, self.consts = set_constants(cfg)
        m = self.metrics
        d = MhdDesc()
        d.nPx, d.nPy, d.nPz = cfg.nPx, cfg.nPy, cfg.nPz
        d.flag_lRot, d.flag_lMag, d.flag_lPot = int(cfg.lRot), int(cfg.lMag), int(cfg.lPot)
        d.flag_lRem, d.flag_lShr = int(cfg.lRem), int(cfg.lShr)
        d.flag_ixCon, d.flag_iyCon, d.flag_izCon = cfg.ixCon, cfg.iyCon, cfg.izCon
        d.flag_iTCon, d.flag_iBCon, d.flag_iD = cfg.iTCon, cfg.iBCon, cfg.iD
        d.precision = cfg.precision
        d.relax_nPEz = cfg.relax_nPEz
        for k, v in self.consts.items():
            setattr(d, k, v)
        self._arrays = {}
        for name in ("dxxdx", "d2xxdx2", "DDX", "dyydy", "d2yydy2", "DDY",
                     "dzzdz", "d2zzdz2", "DDZ", "RKAPA", "DKAPA", "EXX", "WYY", "ZEE"):
            a = np.ascontiguousarray(getattr(m, name), dtype=np.float64)
            self._arrays[name] = a
            setattr(d, name, a.ctypes.data_as(C.POINTER(C.c_double)))
        self.struct = d

    @property
    def shape(self):
        c = self.cfg
        return (c.nPz + 2, c.nPy + 2, c.nPx + 2)

    def ptr(self):
        return C.byref(self.struct)


def set_constants(cfg: RunConfig):
    """Restatement of ``set()`` (3dmhd.cpp:353-495) for one domain."""
    par = np.asarray(cfg.par, dtype=np.float64)
    nPz_g = cfg.nPz_global or cfg.nPz
    nPy_g = cfg.nPy_global or cfg.nPy
    c = {}
    c["C13"] = fdiv(1.0, 3.0)                                   # :354-356
    c["C23"] = 2.0 * c["C13"]
    c["C43"] = 4.0 * c["C13"]
    c["GAM1"] = fdiv(8.0, 15.0)                                 # :358-362
    c["GAM2"] = fdiv(5.0, 12.0)
    c["GAM3"] = fdiv(3.0, 4.0)
    c["ZETA1"] = fdiv(-17.0, 60.0)
    c["ZETA2"] = fdiv(-5.0, 12.0)
    c["SFF"] = f32(0.8)                                         # 3dmhdparam.h:111
    c["hx"] = float(F32(0.5) * F32(cfg.nPx - 1))                # :364-369
    c["h2x"] = float(F32(cfg.nPx - 1) * F32(cfg.nPx - 1))
    c["hy"] = float(F32(0.5) * F32(nPy_g - 1))
    c["h2y"] = float(F32(nPy_g - 1) * F32(nPy_g - 1))
    c["hz"] = float(F32(0.5) * F32(nPz_g - 1))
    c["h2z"] = float(F32(nPz_g - 1) * F32(nPz_g - 1))
    c["ORZF"] = fdiv(1.0, float(nPz_g - 1))                     # 3dmhdsub.h:797,849
    RE, PR, THETA, GRAV = par[0], par[1], par[2], par[3]        # :401-431
    RY, ANG, RM, BETA = par[4], par[5], par[6], par[7]
    PZP, SIGMA, POLYS, TP, TC = par[8], par[9], par[10], par[11], par[12]
    xMax, yMax, zMax = par[16], par[17], par[18]
    GAMMA = par[53]
    c["RE"] = RE
    c["REPR"] = RE * PR                                         # :451-459
    c["ORE"] = 1.0 / RE
    with np.errstate(divide="ignore"):
        c["ORM"] = float(np.float64(1.0) / np.float64(RM))
        c["OBETA"] = float(np.float64(2.0) / np.float64(BETA))
    c["RM"] = RM
    c["GAMMA"] = GAMMA
    c["CV"] = 1.0 / (GAMMA - 1.0)                               # :144-145
    c["OCV"] = 1.0 / c["CV"]
    c["THETA"] = THETA
    c["GRAV"] = GRAV
    if cfg.lRot:                                                # :432-438
        c["OMX"] = math.sin(ANG) / RY
        c["OMZ"] = math.cos(ANG) / RY
    else:
        c["OMX"] = 0.0
        c["OMZ"] = 0.0
    c["TP"] = TP
    c["TB"] = float(cfg.TB)
    c["TC"] = TC
    c["RLAX"] = par[59]                                         # :428-430
    c["TSTART"], c["TOFF"] = par[57], par[58]
    c["TIMI"] = float(cfg.TIMI)
    c["DH"], c["DV"] = par[55], par[56]                        # :449
    c["zMax"] = zMax
    c["TU"], c["DZTU"], c["DZTB"] = float(cfg.TU), float(cfg.DZTU), float(cfg.DZTB)
    c["XP"], c["YP"], c["ZP"] = par[13], par[14], par[15]      # :414-416
    c["QFH"], c["HH"] = par[52], par[54]                       # :424,426
    c["ORZD"] = 1.0 / float(nPz_g - 1)                         # 3dmhdsub.h:768

    if cfg.nPx > 1:                                             # :461-468
        EXX, dxxdx, d2xxdx2, DDX = _uniform_axis(cfg.nPx, cfg.nPx, xMax, ghosts_zero=True)
    else:
        EXX, dxxdx, d2xxdx2, DDX = (np.zeros(cfg.nPx + 2) for _ in range(4))
    WYY, dyydy, d2yydy2, DDY = _uniform_axis(cfg.nPy, nPy_g, yMax, ghosts_zero=False)
    if cfg.zgrid[0] == "uniform":
        ZEE, dzzdz, d2zzdz2, DDZ = _uniform_axis(cfg.nPz, nPz_g, zMax, ghosts_zero=False)
    elif cfg.zgrid[0] == "tanh":
        ZEE, dzzdz, d2zzdz2, DDZ = tanh_axis(cfg.nPz, zMax, float(cfg.zgrid[1]))
    elif cfg.zgrid[0] == "arrays":
        ZEE, dzzdz, d2zzdz2, DDZ = (np.array(a, dtype=np.float64) for a in cfg.zgrid[1:5])
        if any(a.shape != (cfg.nPz + 2,) for a in (ZEE, dzzdz, d2zzdz2, DDZ)):
            raise ValueError("z metric arrays must have nPz + 2 entries")
    else:
        raise ValueError(f"unknown z grid {cfg.zgrid!r}")
    nz = cfg.nPz + 2
    if PZP == 0.0:                                              # :477-488
        RKAPA = np.ones(nz)
        DKAPA = np.zeros(nz)
    else:
        RKAPST = 1.0 if GRAV == 0.0 else (POLYS + 1.0) * THETA / GRAV
        RKAPA = np.array([1.0 + (RKAPST - 1.0) / 2.0 * (1.0 + math.tanh((ZEE[i] - PZP) / SIGMA))
                          for i in range(nz)])
        DKAPA = np.array([(RKAPST - 1.0) / 2.0 / SIGMA / math.cosh((ZEE[i] - PZP) / SIGMA) ** 2
                          for i in range(nz)])
    if cfg.kappa is not None:
        RKAPA = np.asarray(cfg.kappa[0], dtype=np.float64).copy()
        DKAPA = np.asarray(cfg.kappa[1], dtype=np.float64).copy()
    m = Metrics(EXX, dxxdx, d2xxdx2, DDX, WYY, dyydy, d2yydy2, DDY,
                ZEE, dzzdz, d2zzdz2, DDZ, RKAPA, DKAPA)
    return m, c
