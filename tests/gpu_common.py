"""Shared helpers of the GPU parity tests (they compare libmhd3d.so with the oracle)."""
import numpy as np

from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.initial import stratified_config, stratified_state
from oracle import oracle as om

CASES = {
    "hydro": dict(),
    "rot": dict(lRot=True),
    "mhd": dict(lMag=True),
    "mhd_rot": dict(lMag=True, lRot=True),
}

INTERIOR = (slice(1, -1), slice(1, -1), slice(1, -1))   # interior incl. both walls


def relerr(g, c):
    """max|g-c| / max|c| over the interior incl. walls (0/0 -> 0)."""
    g = g[INTERIOR]
    c = c[INTERIOR]
    num = float(np.max(np.abs(g - c)))
    den = float(np.max(np.abs(c)))
    return num / den if den > 0 else num


def bit_identical_fraction(g, c):
    g = g[INTERIOR]
    c = c[INTERIOR]
    return float(np.mean(g.view(np.int64) == c.view(np.int64)))


def make_pair(n, case, seed=7, **kw):
    cfg = stratified_config(*n, **CASES[case], **kw)
    desc = Descriptor(cfg)
    st = stratified_state(desc, seed=seed)
    dom = Domain(desc)
    dom.set_state(st)
    orc = om.Oracle(desc)
    orc.set_state(st)
    orc.communicate()
    return desc, dom, orc, st
