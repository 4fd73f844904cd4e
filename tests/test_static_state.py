"""The product's Static() (mhd3d_static_state, host code) against the oracle's
restatement of the same routine, which the reference's published iteration-9
answers pin (tests/test_oracle_known_answers.py): bit-identical interiors,
including the default 4x6-rank run and a run whose planes do not divide
evenly among the z ranks."""
import os

import numpy as np
import pytest

from known_answers import default_config, default_state
from mhd3d import _lib as L
from mhd3d.initial import reference_static_state
from oracle import oracle as om

pytestmark = pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="libmhd3d.so not built")

INNER = (slice(None), slice(1, -1), slice(1, -1))


def _check(cfg, NP, NPE_Y, NPE_Z):
    ref = default_state(cfg, om, NP=NP, NPE_Y=NPE_Y, NPE_Z=NPE_Z)
    got = reference_static_state(cfg, NP, NPE_Y, NPE_Z)
    for k in ref:
        assert np.array_equal(got[k][INNER], ref[k][INNER]), k
    # x / y ghosts are periodic copies
    TT = got["TT"]
    assert np.array_equal(TT[:, 1:-1, 0], TT[:, 1:-1, -2]) and np.array_equal(TT[:, 0, :], TT[:, -2, :])


def test_default_rank_layout():
    cfg = default_config()          # 256 x 256 x 252 of the 4 x 6 rank run (256 global planes)
    _check(cfg, 256, 4, 6)


def test_small_uneven_layout():
    from dataclasses import replace
    cfg = replace(default_config(), nPx=16, nPy=12, nPz=18)   # 20 planes on 3 z ranks of 6
    _check(cfg, 20, 2, 3)


def test_magnetic_layers_refused():
    from dataclasses import replace
    cfg = replace(default_config(), nPx=16, nPy=12, nPz=18, lMag=True)
    with pytest.raises(L.MhdError) as ei:
        reference_static_state(cfg, 20, 2, 3)
    assert ei.value.code == L.EUNSUPPORTED
