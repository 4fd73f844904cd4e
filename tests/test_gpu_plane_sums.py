"""The plane sums of the FP64 form (horizontal_mean, 3dmhdsub.h:698-718) on
adversarial planes, bit for bit against the oracle's one sequential chain.

The planes make the rounding of the running sum s = s + v hard to reproduce
by anything but the reference's order: values on coarse dyadic grids (exact
halfway cases, both signs, running sums crossing zero and binades), pairs
that cancel exactly to zero, magnitudes over 18 decades, positive planes with
exact halfway cases, planes close to one value (as RO and TT are), and
planes of one value whose running sum passes a binade where every summand
is a halfway case, on planes of up to 512 x 512 summands.  The sums are seen through fluxes() of iD 1 (RO
means), lRem (TT means) and RLAX (TT means and the convective-flux means,
3dmhdsub.h:349-353), compared with the oracle bit for bit.
"""
import numpy as np
import pytest

import test_gpu_ext as tx
import test_gpu_relax as tr
from gpu_common import bit_identical_fraction, relerr
from mhd3d import _lib as L
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.initial import stratified_state
from oracle import oracle as om

pytestmark = pytest.mark.gpu


def _adversarial(st, kind, seed):
    r = np.random.default_rng(seed)
    ro, tt = st["RO"], st["TT"]
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    shape = ro[inner].shape
    if kind == "dyadic":
        ro[inner] = r.integers(8, 64, shape) / 16.0          # 0.5 .. 4 in steps of 1/16
        tt[inner] = r.integers(-32, 33, shape) / 8.0         # -4 .. 4 in steps of 1/8
    elif kind == "cancel":
        v = r.uniform(1.0, 2.0, shape)
        v[..., 1::2] = -v[..., 0::2][..., : v[..., 1::2].shape[-1]]   # a, -a, b, -b, ... along x
        tt[inner] = v
    elif kind == "wide":
        tt[inner] = r.uniform(-1.0, 1.0, shape) * 10.0 ** r.integers(-12, 6, shape)
    elif kind == "halves":
        # positive; every 50th value ends at bit 2^-m, m in 34..46, which is
        # half the grid of a running sum of 2^(53-m) and a halfway case there
        for a in (ro, tt):
            v = r.uniform(1.0, 2.0, shape).reshape(-1)
            sel = v[::50]
            v[::50] = 1.5 + (2 * r.integers(0, 512, sel.shape) + 1) * 2.0 ** -r.integers(34, 47, sel.shape)
            a[inner] = v.reshape(shape)
    elif kind == "const":
        # one value per plane, 1 + odd * 2^-m: every summand is a halfway case
        # of the running sum's grid while that sum is in the binade 2^(53-m)
        for a in (ro, tt):
            for z in range(1, a.shape[0] - 1):
                m = int(r.integers(38, 47))
                a[z, 1:-1, 1:-1] = 1.0 + (2 * int(r.integers(0, 4096)) + 1) * 2.0 ** -m
    else:   # temp: positive, close to one value, as temperatures and densities are
        ro[inner] = 1.0 + 1e-3 * r.standard_normal(shape)
        tt[inner] = 2.0 + 1e-3 * r.standard_normal(shape)
    return st


def _check(cfg, kind, seed):
    desc = Descriptor(cfg)
    st = _adversarial(stratified_state(desc, seed=seed), kind, seed)
    dom = Domain(desc)
    dom.set_state(st)
    orc = om.Oracle(desc)
    orc.set_state(st)
    orc.communicate()
    fg = dom.fluxes()
    orc.fluxes()
    for name in L.FLUXES[: dom.nf]:
        c = orc.get(name)
        assert bit_identical_fraction(fg[name], c) == 1.0, (kind, name, relerr(fg[name], c))
    dom.close()
    orc.close()


@pytest.mark.parametrize("case,n,kind", [("iD1", (40, 30, 10), "dyadic"), ("lRem1", (40, 30, 10), "dyadic"),
                                         ("lRem1", (40, 30, 10), "cancel"), ("lRem1", (40, 30, 10), "wide"),
                                         ("lRem1", (128, 64, 8), "dyadic"), ("iD1", (3, 200, 8), "dyadic")])
def test_branch_sums_on_adversarial_planes(case, n, kind):
    print(f"plane sums {case} {n} {kind}", flush=True)
    _check(tx._cfg(n, case), kind, 21)


# planes of many chunks of 2048 summands, whole and partial
@pytest.mark.parametrize("case,n,kind", [("lRem1", (256, 96, 6), "dyadic"), ("lRem1", (250, 90, 6), "wide"),
                                         ("lRem1", (250, 90, 6), "cancel"), ("iD1", (200, 100, 6), "halves"),
                                         ("lRem1", (200, 100, 6), "halves"), ("iD1", (333, 77, 6), "temp"),
                                         ("lRem1", (512, 512, 5), "temp"), ("iD1", (512, 512, 5), "dyadic"),
                                         ("iD1", (512, 512, 8), "const"), ("lRem1", (300, 200, 8), "const")])
def test_branch_sums_on_large_planes(case, n, kind):
    print(f"plane sums {case} {n} {kind}", flush=True)
    _check(tx._cfg(n, case), kind, 29)


@pytest.mark.parametrize("kind", ["dyadic", "cancel", "wide"])
def test_relaxation_sums_on_adversarial_planes(kind):
    print(f"plane sums rlax {kind}", flush=True)
    _check(tr._cfg((40, 30, 10), "it0_mhd", P=2), kind, 23)
