"""Synthetic stratified-convection runs (the workloads the library is measured on).

The reference has no Rayleigh/Taylor/stretch parameters: its runs are set by
``par[64]`` (3dmhdset.h:7-24, meaning documented at :239-303).  The mapping
used here, fixed once and used by tests and benchmarks alike:

* polytrope of index m with density contrast D top-to-bottom: T = 1 + THETA*z,
  rho = T**m (the hydrostatic solution of ``derivs``, 3dmhdphsic.h:347-355,
  with GRAV = (m+1)*THETA), THETA*zMax = D**(1/m) - 1, zMax = 1;
* Rayleigh number at the top (rho = T = 1), depth 1:
  Ra = (m+1) THETA^2 (1 - (m+1)(GAMMA-1)/GAMMA) RE^2 PR, solved for RE;
* Taylor number Ta = (RE/RY)^2 (2*Omega = 1/RY in the Coriolis terms
  OMX = sin(ANG)/RY, OMZ = cos(ANG)/RY, 3dmhd.cpp:432-434), ANG = 0.25;
* magnetic Prandtl number Pm = RM/RE; plasma beta BETA = 1 (OBETA = 2);
* walls: fixed temperature top (iTCon = 0, TP = 0 -> T = 1) and bottom
  (izCon = 0, TB = 1 + THETA*zMax), stress-free and impenetrable (BCON);
* x, y uniform and periodic with xMax = yMax = 3*zMax; z tanh-stretched
  (clustered at both walls) with stretch factor beta;
* perturbations: RU, RV, RW = rho * amp * U(-1, 1) on the planes between the
  walls, from numpy's default_rng(seed); with lMag, BX = B0 + b_amp*U(-1,1),
  BY, BZ = b_amp*U(-1,1), BZ = 0 on the walls.
"""
from __future__ import annotations

import math

import numpy as np

from .desc import Descriptor, RunConfig, default_par, fdiv


def stratified_config(nPx, nPy, nPz, Ra=1e5, Pr=1.0, lMag=False, lRot=False, Ta=1e4,
                      Pm=1.0, contrast=10.0, m_poly=1.0, stretch=1.5, aspect=3.0,
                      beta=1.0, ang=0.25, precision=0):
    par = default_par()
    gamma = fdiv(5.0, 3.0)
    zmax = 1.0
    theta = (contrast ** (1.0 / m_poly) - 1.0) / zmax
    grav = (m_poly + 1.0) * theta
    coeff = (m_poly + 1.0) * theta ** 2 * (1.0 - (m_poly + 1.0) * (gamma - 1.0) / gamma)
    re = math.sqrt(Ra / (coeff * Pr))
    par[0] = re
    par[1] = Pr
    par[2] = theta
    par[3] = grav
    par[4] = re / math.sqrt(Ta) if lRot else 0.0
    par[5] = ang if lRot else 0.0
    par[6] = Pm * re if lMag else 0.0
    par[7] = beta if lMag else 0.0
    par[11] = 0.0
    par[16] = aspect * zmax
    par[17] = aspect * zmax
    par[18] = zmax
    par[53] = gamma
    return RunConfig(nPx=nPx, nPy=nPy, nPz=nPz, lRot=lRot, lMag=lMag, izCon=0, iTCon=0,
                     par=par, TB=1.0 + theta * zmax, zgrid=("tanh", stretch),
                     precision=precision)


def periodic_fill(a):
    """Refill x and y ghost cells periodically (communicate(), 3dmhdmpi.h:59-142)."""
    a[:, 1:-1, 0] = a[:, 1:-1, -2]
    a[:, 1:-1, -1] = a[:, 1:-1, 1]
    a[:, 0, :] = a[:, -2, :]
    a[:, -1, :] = a[:, 1, :]
    return a


def stratified_profiles(desc: Descriptor):
    """z profiles (TT_z, RO_z) of the polytrope, ghost planes included, with
    the wall temperatures BCON imposes."""
    cfg = desc.cfg
    par = cfg.par
    theta, grav = par[2], par[3]
    m_poly = grav / theta - 1.0
    zee = np.array(desc.metrics.ZEE, dtype=np.float64)
    T = 1.0 + theta * zee
    rho = T ** m_poly
    T[-1] = T[-2] = cfg.TB if cfg.izCon == 0 else T[-2]
    T[0] = T[1] = 1.0
    return T, rho


def stratified_state(desc: Descriptor, seed=12345, amp=1e-4, b0=1e-3, b_amp=1e-5):
    cfg = desc.cfg
    nz, ny, nx = desc.shape
    T, rho = stratified_profiles(desc)
    TT = np.broadcast_to(T[:, None, None], (nz, ny, nx)).copy()
    RO = np.broadcast_to(rho[:, None, None], (nz, ny, nx)).copy()
    rng = np.random.default_rng(seed)
    inner = (slice(2, nz - 2), slice(1, ny - 1), slice(1, nx - 1))
    out = {}
    for name in ("RU", "RV", "RW"):
        a = np.zeros((nz, ny, nx))
        a[inner] = RO[inner] * amp * rng.uniform(-1.0, 1.0, size=a[inner].shape)
        out[name] = periodic_fill(a)
    out["RO"] = periodic_fill(RO)
    out["TT"] = periodic_fill(TT)
    if cfg.lMag:
        walls = (slice(1, nz - 1), slice(1, ny - 1), slice(1, nx - 1))
        BX = np.zeros((nz, ny, nx))
        BY = np.zeros((nz, ny, nx))
        BZ = np.zeros((nz, ny, nx))
        BX[walls] = b0 + b_amp * rng.uniform(-1.0, 1.0, size=BX[walls].shape)
        BY[walls] = b_amp * rng.uniform(-1.0, 1.0, size=BY[walls].shape)
        BZ[walls] = b_amp * rng.uniform(-1.0, 1.0, size=BZ[walls].shape)
        BZ[1] = 0.0
        BZ[-2] = 0.0
        for B in (BX, BY, BZ):       # finite ghost planes: copies of the walls
            B[0] = B[1]
            B[-1] = B[-2]
        out["BX"] = periodic_fill(BX)
        out["BY"] = periodic_fill(BY)
        out["BZ"] = periodic_fill(BZ)
    return out


_GOLD = np.uint64(0x9E3779B97F4A7C15)


def hash_uniform(seed, f, z, y, x):
    """Counter-based uniform deviate in [-1, 1): the splitmix64 finaliser of
    the packed key (the device initializer's generator, bit for bit)."""
    z = np.asarray(z, dtype=np.uint64)
    y = np.asarray(y, dtype=np.uint64)
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        k = (np.uint64(seed) * _GOLD) ^ (np.uint64(f) << np.uint64(56)) ^ (z << np.uint64(36)) \
            ^ (y << np.uint64(18)) ^ x
        k = k + _GOLD
        k = (k ^ (k >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        k = (k ^ (k >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        k = k ^ (k >> np.uint64(31))
    return (k >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def stratified_state_hashed(desc: Descriptor, seed=12345, amp=1e-4, b0=1e-3, b_amp=1e-5):
    """Host copy of the state mhd3d_init_stratified builds on the device."""
    cfg = desc.cfg
    nz, ny, nx = desc.shape
    T, rho = stratified_profiles(desc)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny - 2), np.arange(nx - 2), indexing="ij")
    out = {}
    RO = np.zeros((nz, ny, nx))
    RO[:, 1:-1, 1:-1] = rho[:, None, None]
    TT = np.zeros((nz, ny, nx))
    TT[:, 1:-1, 1:-1] = T[:, None, None]
    inner = (zz >= 2) & (zz <= nz - 3)
    for f, name in enumerate(("RU", "RV", "RW")):
        a = np.zeros((nz, ny, nx))
        a[:, 1:-1, 1:-1] = np.where(inner, RO[:, 1:-1, 1:-1] * amp * hash_uniform(seed, f, zz, yy, xx), 0.0)
        out[name] = periodic_fill(a)
    out["RO"] = periodic_fill(RO)
    out["TT"] = periodic_fill(TT)
    if cfg.lMag:
        zw = np.where(zz == 0, 1, np.where(zz == nz - 1, nz - 2, zz))
        wall = (zw == 1) | (zw == nz - 2)
        BX = np.zeros((nz, ny, nx))
        BY = np.zeros((nz, ny, nx))
        BZ = np.zeros((nz, ny, nx))
        BX[:, 1:-1, 1:-1] = b0 + b_amp * hash_uniform(seed, 5, zw, yy, xx)
        BY[:, 1:-1, 1:-1] = b_amp * hash_uniform(seed, 6, zw, yy, xx)
        BZ[:, 1:-1, 1:-1] = np.where(wall, 0.0, b_amp * hash_uniform(seed, 7, zw, yy, xx))
        out["BX"] = periodic_fill(BX)
        out["BY"] = periodic_fill(BY)
        out["BZ"] = periodic_fill(BZ)
    return out


def reference_static_state(cfg, nPz_global, nPEy=1, nPEz=1, with_profiles=False):
    """The reference's Static() initial state (mhd3d_static_state_flags): the
    run of nPEy x nPEz ranks on a grid of nPz_global planes, assembled into one
    domain of nPEz*(nPz_global//nPEz) planes (cfg.nPz must be that), with the
    lMag, lShr and lRem switches of cfg.  with_profiles: also return what
    Static() sets besides the arrays under lRem, {"RKAPA", "DKAPA", "TU",
    "DZTU", "DZTB"} (the descriptor's kappa and wall temperatures)."""
    import ctypes as C
    from . import _lib as L

    nRz = nPz_global // nPEz
    if cfg.nPz != nPEz * nRz:
        raise ValueError(f"cfg.nPz must be {nPEz * nRz} for {nPEz} z ranks of {nPz_global} planes")
    shape = (cfg.nPz + 2, cfg.nPy + 2, cfg.nPx + 2)
    nf = 8 if cfg.lMag else 5
    arrs = [np.zeros(shape) for _ in range(nf)]
    par = np.ascontiguousarray(cfg.par, dtype=np.float64)
    rk = np.zeros(shape[0])
    drk = np.zeros(shape[0])
    walls = np.zeros(3)
    dp = C.POINTER(C.c_double)
    rc = L.lib().mhd3d_static_state_flags(cfg.nPx, cfg.nPy, int(nPz_global), int(nPEy), int(nPEz),
                                          par.ctypes.data_as(dp), int(cfg.lMag), int(cfg.lShr),
                                          int(cfg.lRem), L._fieldptrs(arrs), rk.ctypes.data_as(dp),
                                          drk.ctypes.data_as(dp), walls.ctypes.data_as(dp))
    if rc != L.OK:
        raise L.MhdError(rc, "mhd3d_static_state_flags failed")
    st = dict(zip(L.FIELDS[:nf], arrs))
    if not with_profiles:
        return st
    prof = {"RKAPA": rk, "DKAPA": drk, "TU": walls[0], "DZTU": walls[1], "DZTB": walls[2]}
    return st, prof
