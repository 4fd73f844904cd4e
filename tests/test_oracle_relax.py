"""The oracle's RLAX heat-flux relaxation against what the mathematics fixes (CPU).

No published data covers RLAX (3dmhdsub.h:322-442), so each part is pinned by
a closed form of the same quantity, written here without the oracle's loops:

* the convective flux mean FCONM: for RW = B0 + B cos(kx), TT = t0(z) +
  A cos(kx), constant density rho and no horizontal motion, the plane mean of
  -2.5 RW T' - 0.5 RW w^2 (T' = TT - TTM, w = RW / rho) is
  -1.25 A B - 0.5 (B0^3 + 1.5 B0 B^2) / rho^2 (the discrete means of cos,
  cos^3 vanish and that of cos^2 is 1/2 on 16 points);
* the radiative flux FRADM: for a temperature linear in z, slope * RKAPA / REPR
  on every plane (the central and the one-sided wall differences are exact);
* CORRECT: with a constant ALPHA the per-rank trapezoid scans add up to
  alpha' (ZEE - ZEE[1]), alpha' = alpha / (1 + 4|alpha|); where FCONM != 0
  each rank boundary below a plane adds 0.5 (b' - a') dz (the ghost plane of
  a rank takes FCONM = 0); with iTCon 1 / izCon 0 every rank subtracts twice
  the end values of itself and all ranks above it, the last one's ghost
  segment with ALPHA = 1 / 5;
* FT: the relaxation adds CORRECT(z) * RLAX * TFAC to FT and nothing else;
* TFAC: +DT / (3 TOFF) on a new extreme of the wall-plane mean (capped at 2),
  -DT / (3 TOFF) otherwise, frozen once at or below 1e-4f, untouched before
  TSTART;
* the exit points: a fluxes() call that opens a step (ICALL % 3 == 0) with
  fmod(TIMT, 25) <= DT stops the reference's program (iTCon 0 / izCon 1 and
  iTCon 1 / izCon 0), as does any active relaxation with other walls.

Planted mistakes 23-27 (TT in place of T', no 1/(1+4|a|), the real FCONM
on the rank ghosts, one subtraction in the iTCon 1 sum, DT/TOFF steps of
TFAC) are each caught (tests/test_oracle_mutants.py).
Tolerances: 1e-12 relative for the closed forms (roundings of sums of a few
hundred terms), exact for the TFAC/ICALL bookkeeping.
"""
import math

import numpy as np
import pytest

from mhd3d.desc import Descriptor, RunConfig, default_par, _uniform_axis
from oracle import oracle as om
from test_oracle_math import _coords

NPX, NPY, NPZ = 16, 12, 12


def _par(RLAX=0.3, TSTART=1.0, TOFF=0.5, THETA=3.0):
    par = default_par()
    par[0] = 40.0           # RE
    par[1] = 0.7            # PR
    par[2] = THETA
    par[3] = 8.0            # GRAV
    par[16], par[17], par[18] = 2.0, 1.5, 1.0
    par[57], par[58], par[59] = TSTART, TOFF, RLAX
    return par


def _zgrid(ghost_dz=True):
    """The uniform z axis, with the ghost coordinates continued linearly
    (so the last rank's ghost segment has a length)."""
    ZEE, dzz, d2zz, DDZ = _uniform_axis(NPZ, NPZ, 1.0, ghosts_zero=False)
    ZEE = np.array(ZEE, dtype=np.float64)
    if ghost_dz:
        ZEE[0] = 2 * ZEE[1] - ZEE[2]
        ZEE[-1] = 2 * ZEE[-2] - ZEE[-3]
    return ("arrays", ZEE, dzz, d2zz, DDZ)


def _desc(P=2, iTCon=0, izCon=1, TIMI=10.5, lRem=False, **kw):
    cfg = RunConfig(nPx=NPX, nPy=NPY, nPz=NPZ, iTCon=iTCon, izCon=izCon, par=_par(**kw),
                    zgrid=_zgrid(), TIMI=TIMI, relax_nPEz=P, lRem=lRem)
    return Descriptor(cfg)


def _cosx(desc):
    nz, ny, nx = desc.shape
    th = 2 * np.pi * (np.arange(nx) - 1) / NPX
    return np.broadcast_to(np.cos(th), (nz, ny, nx)).copy()


def _state(desc, slope=-0.8, A=0.0, B=0.0, B0=0.0, rho=1.3):
    _, _, Z = _coords(desc)
    TT = 1.0 + slope * Z + A * _cosx(desc)
    RW = B0 + B * _cosx(desc)
    z = np.zeros(desc.shape)
    return {"RU": z, "RV": z.copy(), "RW": RW, "RO": np.full(desc.shape, rho), "TT": TT}


def _run(desc, st, dt=1e-3, timt=None, calls=1):
    o = om.Oracle(desc)
    o.set_state(st)
    o.communicate()
    o.set_time(dt, desc.cfg.TIMI if timt is None else timt)
    for _ in range(calls):
        o.fluxes()
    return o


def _ftot(desc):
    c = desc.consts
    if desc.cfg.lRem:
        return float(np.float32(np.float32(8.07) * np.float32(0.4))) * c["THETA"]
    return c["THETA"] / c["REPR"]


def _lim(a):
    return a / (1.0 + 4.0 * abs(a))


def _slope(desc, st):
    """The z slope of TT in the stencil's coordinate (D1z of a linear field)."""
    _, _, Z = _coords(desc)
    return (st["TT"][2, 1, 1] - st["TT"][1, 1, 1]) / (Z[2, 1, 1] - Z[1, 1, 1])


def _close(a, b, tol):
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= tol * scale, (np.abs(a - b).max(), scale)


# ---------------------------------------------------------------- FRADM, CORRECT


@pytest.mark.parametrize("P", [2, 3, 4])
def test_correct_is_alpha_times_height_at_rest(P):
    desc = _desc(P=P)
    st = _state(desc, slope=-0.8)
    o = _run(desc, st)
    rs = o.relax_state()
    c, m = desc.consts, desc.metrics
    q = _slope(desc, st) * m.RKAPA[1] / c["REPR"]
    assert np.allclose(m.RKAPA, m.RKAPA[1])
    _close(rs["FRADM"][1:-1], np.full(NPZ, q), 1e-12)
    assert np.abs(rs["FCONM"][1:-1]).max() == 0.0           # RW = 0: no convective flux
    ap = _lim(1.0 - q / _ftot(desc))
    ZEE = m.ZEE
    _close(rs["CORRECT"][1:-1], ap * (ZEE[1:-1] - ZEE[1]), 1e-12)
    # the FT term: CORRECT * RLAX * TFAC, with TFAC = 1 + DT/(3 TOFF) after the first call
    tfac = 1.0 + 1e-3 / (3.0 * 0.5)
    assert rs["TFAC"] == tfac and rs["ICALL"] == 1 and rs["exited"] == 0
    off = _run(_desc(P=P, RLAX=0.0), st)
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    dFT = (o.get("FT") - off.get("FT"))[inner]
    exp = np.broadcast_to((rs["CORRECT"] * (0.3 * tfac))[:, None, None], desc.shape)[inner]
    scale = max(np.abs(exp).max(), np.abs(off.get("FT")[inner]).max())
    assert np.abs(dFT - exp).max() <= 1e-12 * scale
    for name in ("FU", "FV", "FW", "FR"):
        assert np.array_equal(o.get(name), off.get(name)), name


def test_convective_flux_mean():
    desc = _desc(P=2)
    A, B, B0, rho = 0.07, 0.05, 0.02, 1.3
    st = _state(desc, slope=-0.8, A=A, B=B, B0=B0, rho=rho)
    rs = _run(desc, st).relax_state()
    exp = -1.25 * A * B - 0.5 * (B0 ** 3 + 1.5 * B0 * B ** 2) / rho ** 2
    _close(rs["FCONM"][1:-1], np.full(NPZ, exp), 1e-12)
    assert rs["FCONM"][0] == 0.0 and rs["FCONM"][-1] == 0.0


@pytest.mark.parametrize("P", [2, 3])
def test_rank_boundaries_take_no_convective_flux(P):
    """FCONM = F on every plane but the ranks' ghost planes (0 there)."""
    desc = _desc(P=P)
    A, B, rho = 0.3, 0.2, 1.3
    st = _state(desc, slope=-0.8, A=A, B=B, rho=rho)
    rs = _run(desc, st).relax_state()
    c, m = desc.consts, desc.metrics
    F = -1.25 * A * B
    q = _slope(desc, st) * m.RKAPA[1] / c["REPR"]
    ftot = _ftot(desc)
    a, b = _lim(1.0 - (F + q) / ftot), _lim(1.0 - q / ftot)
    nRz = NPZ // P
    dz = m.ZEE[2] - m.ZEE[1]
    g = np.arange(1, NPZ + 1)
    r = (g - 1) // nRz
    exp = a * (m.ZEE[g] - m.ZEE[1]) + r * 0.5 * (b - a) * dz
    _close(rs["CORRECT"][1:-1], exp, 1e-12)


@pytest.mark.parametrize("P", [2, 3])
def test_itcon1_subtracts_each_rank_end_twice(P):
    desc = _desc(P=P, iTCon=1, izCon=0)
    A, B, rho = 0.3, 0.2, 1.3
    st = _state(desc, slope=-0.8, A=A, B=B, rho=rho)
    rs = _run(desc, st).relax_state()
    c, m = desc.consts, desc.metrics
    q = _slope(desc, st) * m.RKAPA[1] / c["REPR"]
    ftot = _ftot(desc)
    a, b = _lim(1.0 - (-1.25 * A * B + q) / ftot), _lim(1.0 - q / ftot)
    nRz = NPZ // P
    dz = m.ZEE[2] - m.ZEE[1]
    dzg = m.ZEE[-1] - m.ZEE[-2]               # the last rank's ghost segment
    end = [(nRz - 1) * a * dz + 0.5 * (a + b) * dz for _ in range(P - 1)]
    end.append((nRz - 1) * a * dz + 0.5 * (a + _lim(1.0)) * dzg)
    g = np.arange(1, NPZ + 1)
    r = (g - 1) // nRz
    exp = a * (m.ZEE[g] - m.ZEE[r * nRz + 1]) - 2.0 * np.array([sum(end[k] for k in range(ri, P)) for ri in r])
    _close(rs["CORRECT"][1:-1], exp, 1e-12)
    # TFAC follows the top wall: the first call finds a new minimum of TTM(zl)
    assert rs["TTM_MIN"] == pytest.approx(st["TT"][1, 1:-1, 1:-1].mean(), rel=1e-14)
    assert rs["TFAC"] == 1.0 + 1e-3 / 1.5


def test_lrem_takes_the_rempel_total_flux():
    desc = _desc(P=2, lRem=True)
    st = _state(desc, slope=-0.8)
    rs = _run(desc, st).relax_state()
    c, m = desc.consts, desc.metrics
    q = _slope(desc, st) * m.RKAPA[1] / c["REPR"]
    ftot = float(np.float32(8.07) * np.float32(0.4)) * c["THETA"]
    _close(rs["CORRECT"][1:-1], _lim(1.0 - q / ftot) * (m.ZEE[1:-1] - m.ZEE[1]), 1e-12)


# ---------------------------------------------------------------- TFAC, gates, exits


def test_tfac_steps_and_cap():
    desc = _desc(P=2, TOFF=0.5)
    st = _state(desc)
    s = 0.3 / (3.0 * 0.5)
    o = _run(desc, st, dt=0.3, calls=1)
    assert o.relax_state()["TFAC"] == 1.0 + s
    o.fluxes()                        # same wall mean: no new maximum, TFAC falls
    assert o.relax_state()["TFAC"] == 1.0 + s - s
    o.fluxes()
    assert o.relax_state()["TFAC"] == 1.0 + s - s - s
    assert o.relax_state()["TTM_MAX"] == pytest.approx(st["TT"][NPZ, 1:-1, 1:-1].mean(), rel=1e-14)
    assert o.relax_state()["ICALL"] == 3
    big = _run(desc, st, dt=6.0)      # DT/(3 TOFF) = 4: capped at 2
    assert big.relax_state()["TFAC"] == 2.0


def test_tfac_below_threshold_freezes_the_relaxation():
    desc = _desc(P=2, TOFF=0.5)
    st = _state(desc)
    s = 0.9 / (3.0 * 0.5)
    o = _run(desc, st, dt=0.9, calls=1)          # steps of 0.6: 1.6, 1.0, 0.4, -0.2, frozen
    for _ in range(3):
        o.fluxes()
    tf = o.relax_state()["TFAC"]
    assert tf == 1.0 + s - s - s - s
    off = _run(_desc(P=2, RLAX=0.0), st)
    o.fluxes()
    rs = o.relax_state()
    assert rs["TFAC"] == tf and rs["ICALL"] == 5
    assert np.all(rs["CORRECT"][1:-1] == 0.0)
    assert np.array_equal(o.get("FT")[1:-1, 1:-1, 1:-1], off.get("FT")[1:-1, 1:-1, 1:-1])


def test_before_tstart_nothing_changes():
    desc = _desc(P=2, TSTART=20.0)
    st = _state(desc, A=0.3, B=0.2)
    o = _run(desc, st)
    off = _run(_desc(P=2, RLAX=0.0, TSTART=20.0), st)
    rs = o.relax_state()
    assert rs["TFAC"] == 1.0 and rs["TTM_MAX"] == -1e99 and rs["ICALL"] == 1
    assert np.all(rs["CORRECT"][1:-1] == 0.0)
    for name in ("FU", "FV", "FW", "FR", "FT"):
        assert np.array_equal(o.get(name)[1:-1, 1:-1, 1:-1], off.get(name)[1:-1, 1:-1, 1:-1]), name


def test_fresh_run_exits_in_its_first_call():
    desc = _desc(P=2, TIMI=0.0, TSTART=0.0)
    o = _run(desc, _state(desc))
    assert o.relax_state()["exited"] == 1


@pytest.mark.parametrize("walls", [(0, 1), (1, 0)])
def test_restart_exits_when_a_step_opens_within_dt_of_25(walls):
    """TIMT = 20.2 + 0.5 k at the start of step k: the first k with
    fmod(TIMT, 25) <= 0.5 is k = 10 (TIMT = 25.2)."""
    desc = _desc(P=2, iTCon=walls[0], izCon=walls[1], TIMI=20.2, TSTART=1.0, TOFF=50.0)
    o = om.Oracle(desc)
    o.set_state(_state(desc))
    o.communicate()
    timc = 0.0
    for k in range(12):
        timt = 20.2 + timc                    # TIMT = TIMI + TIMC (3dmhdsub.h:150-152)
        o.set_time(0.5, timt)
        o.fluxes()                            # the step's three fluxes() calls
        ex = o.relax_state()["exited"]
        assert ex == (1 if k >= 10 else 0), (k, timt, ex)
        if ex:
            assert math.fmod(timt, 25.0) <= 0.5
            break
        o.fluxes()
        o.fluxes()
        timc += 0.5
    assert o.relax_state()["ICALL"] == 31


def test_other_walls_exit_once_the_relaxation_starts():
    st_desc = _desc(P=2, iTCon=4, izCon=1, TIMI=30.0, TSTART=40.0)
    o = _run(st_desc, _state(st_desc), timt=30.0)       # before TSTART: runs, even at 25 k + small
    assert o.relax_state()["exited"] == 0
    o.set_time(1e-3, 50.0)
    o.fluxes()
    assert o.relax_state()["exited"] == 2
