"""Synthetic workloads: the ensemble mix and the served request trace.

Ensemble mix: runs with nx = ny drawn from {32, 48, 64, 96, 128} and nz from
{32, 48, 64, 96}, Rayleigh number log-uniform in [1e4, 1e6], Prandtl and
magnetic Prandtl numbers log-uniform in [0.3, 3], lMag on in half of the runs
(even indices), lRot on at random, per-run perturbation seed base_seed + i.

Served trace: Poisson arrivals (mean rate `rate` per second) of runs with
grids from 32^3 up to 256x256x128, lMag and lRot drawn at random, parameters
as in the ensemble mix and step counts log-uniform in [50, 2000].

Magnetic runs keep a magnetic Reynolds number RM = Pm*RE of at least
`rm_min` (50): Pm is raised to rm_min/RE where the draw falls below.  Below
that, the one-sided magnetic diffusion of the wall planes (3dmhdsub.h:638-690)
drives BX, BY on coarse grids into exponential growth within a few hundred
steps -- in the reference's scheme itself, alone or batched.
"""
from __future__ import annotations

from dataclasses import dataclass

import math

import numpy as np

from .initial import stratified_config

RM_MIN = 50.0


@dataclass
class RunSpec:
    index: int
    nPx: int
    nPy: int
    nPz: int
    Ra: float
    Pr: float
    Pm: float
    lMag: bool
    lRot: bool
    seed: int
    steps: int = 0
    arrival: float = 0.0

    @property
    def cells(self):
        return self.nPx * self.nPy * self.nPz

    def config(self, precision=0):
        return stratified_config(self.nPx, self.nPy, self.nPz, Ra=self.Ra, Pr=self.Pr, Pm=self.Pm,
                                 lMag=self.lMag, lRot=self.lRot, precision=precision)


def _loguniform(rng, lo, hi):
    return float(10.0 ** rng.uniform(np.log10(lo), np.log10(hi)))


def _reynolds(Ra, Pr, contrast=10.0, m_poly=1.0, gamma=5.0 / 3.0):
    """RE of stratified_config for (Ra, Pr)."""
    theta = contrast ** (1.0 / m_poly) - 1.0
    coeff = (m_poly + 1.0) * theta ** 2 * (1.0 - (m_poly + 1.0) * (gamma - 1.0) / gamma)
    return math.sqrt(Ra / (coeff * Pr))


def _pm(Pm, Ra, Pr, lMag, rm_min):
    return max(Pm, rm_min / _reynolds(Ra, Pr)) if lMag else Pm


def ensemble_mix(n=64, seed=2024, base_seed=1000, rm_min=RM_MIN):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nxy = int(rng.choice([32, 48, 64, 96, 128]))
        nz = int(rng.choice([32, 48, 64, 96]))
        Ra, Pr, Pm = _loguniform(rng, 1e4, 1e6), _loguniform(rng, 0.3, 3.0), _loguniform(rng, 0.3, 3.0)
        lMag = i % 2 == 0
        out.append(RunSpec(index=i, nPx=nxy, nPy=nxy, nPz=nz, Ra=Ra, Pr=Pr, Pm=_pm(Pm, Ra, Pr, lMag, rm_min),
                           lMag=lMag, lRot=bool(rng.random() < 0.5), seed=base_seed + i))
    return out


def served_trace(n=200, rate=50.0, seed=7, base_seed=5000,
                 grids=((32, 32, 32), (48, 48, 32), (64, 64, 32), (64, 64, 64), (96, 96, 48),
                        (128, 128, 64), (128, 128, 96), (192, 192, 96), (256, 256, 128)),
                 steps_range=(50, 2000), rm_min=RM_MIN):
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        g = grids[int(rng.integers(len(grids)))]
        steps = int(round(_loguniform(rng, *steps_range)))
        Ra, Pr, Pm = _loguniform(rng, 1e4, 1e6), _loguniform(rng, 0.3, 3.0), _loguniform(rng, 0.3, 3.0)
        lMag = bool(rng.random() < 0.5)
        out.append(RunSpec(index=i, nPx=g[0], nPy=g[1], nPz=g[2], Ra=Ra, Pr=Pr, Pm=_pm(Pm, Ra, Pr, lMag, rm_min),
                           lMag=lMag, lRot=bool(rng.random() < 0.5),
                           seed=base_seed + i, steps=steps, arrival=t))
    return out
