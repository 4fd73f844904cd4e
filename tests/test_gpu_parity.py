"""Parity of the HIP path (through the C ABI) with the oracle on seeded inputs.

Bar: max|g-c|/max|c| <= 1e-12 per field (FP64 plain form); the arithmetic
follows the reference's operation order, so fluxes are also checked for
bit-identity.
"""
import numpy as np
import pytest

from gpu_common import CASES, bit_identical_fraction, make_pair, relerr
from mhd3d import _lib as L

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("case", list(CASES))
def test_fluxes_bit_identical(case):
    desc, dom, orc, _ = make_pair((24, 20, 12), case)
    fg = dom.fluxes()
    orc.fluxes()
    for name in L.FLUXES[: dom.nf]:
        c = orc.get(name)
        frac = bit_identical_fraction(fg[name], c)
        assert frac == 1.0, (case, name, frac, relerr(fg[name], c))


@pytest.mark.parametrize("case", list(CASES))
def test_stages_match(case):
    desc, dom, orc, _ = make_pair((20, 24, 10), case, seed=3)
    dt, _, _ = orc.cfl()
    for k in (1, 2, 3):
        dom.stage(k, dt)
        orc.stage(k, dt)
        sg = dom.state()
        for name, g in sg.items():
            c = orc.get(name)
            assert bit_identical_fraction(g, c) == 1.0, (case, k, name, relerr(g, c))


@pytest.mark.parametrize("case", list(CASES))
def test_cfl_match(case):
    desc, dom, orc, _ = make_pair((16, 18, 9), case, seed=5)
    dt_g, w_g, um_g = dom.cfl()
    dt_c, w_c, um_c = orc.cfl()
    assert dt_g == dt_c
    n = 4 if desc.cfg.lMag else 3
    assert np.array_equal(w_g[:n], w_c[:n])
    assert um_g == um_c


@pytest.mark.parametrize("case", list(CASES))
def test_ten_steps_cfl(case):
    desc, dom, orc, _ = make_pair((32, 24, 16), case, seed=11)
    dts_c = [orc.step() for _ in range(10)]
    dt_g = dom.step(10)
    assert dt_g == dts_c[-1]
    sg = dom.state()
    for name, g in sg.items():
        e = relerr(g, orc.get(name))
        assert e <= TOL, (case, name, e)
    n, timt, timc = dom.clock()
    n_c, timt_c, timc_c = orc.clock()
    assert n == n_c == 10
    assert timc == timc_c


@pytest.mark.parametrize("case", ["hydro", "mhd_rot"])
def test_fixed_dt_and_odd_sizes(case):
    desc, dom, orc, _ = make_pair((33, 17, 7), case, seed=19)
    dt = 1e-4
    for _ in range(3):
        orc.step(fixed_dt=True, dt=dt)
    dom.step(3, dt=dt)
    for name, g in dom.state().items():
        assert bit_identical_fraction(g, orc.get(name)) == 1.0, (case, name)


def test_bcon_match():
    desc, dom, orc, st = make_pair((20, 16, 9), "mhd", seed=23)
    dom.bcon()
    orc.bcon()
    for name, g in dom.state().items():
        assert bit_identical_fraction(g, orc.get(name)) == 1.0, name


def test_two_dimensional_x():
    """nPx = 1: the reference skips xJacobi and drops DDX from the CFL (3dmhdsub.h:77)."""
    desc, dom, orc, _ = make_pair((1, 24, 12), "hydro", seed=29)
    dts_c = [orc.step() for _ in range(3)]
    assert dom.step(3) == dts_c[-1]
    for name, g in dom.state().items():
        assert relerr(g, orc.get(name)) <= TOL, name


def test_unsupported_flags_fail_loudly():
    from mhd3d.desc import Descriptor
    from mhd3d.domain import Domain
    from mhd3d.initial import stratified_config
    cfg = stratified_config(16, 16, 8, lMag=True)
    cfg.iBCon = 1            # only the reference's iBCon 0 wall field is built
    with pytest.raises(L.MhdError) as ei:
        Domain(Descriptor(cfg))
    assert ei.value.code == L.EUNSUPPORTED


@pytest.mark.parametrize("tune", ["1,0", "1,1", "1,2", "1,3", "2,0", "2,1", "2,2", "2,3"])
@pytest.mark.parametrize("case", ["hydro", "mhd_rot", "mhd"])
def test_schedule_variants_bit_identical(case, tune, monkeypatch):
    """Every streaming-kernel variant (register budget, barriers, deferred
    update) of the FP64 form is bit-identical to the oracle."""
    env = "MHD3D_TUNE_FP64_" + ("MHD" if "mhd" in case else "HYDRO")
    monkeypatch.setenv(env, tune)
    desc, dom, orc, _ = make_pair((40, 24, 12), case, seed=31)
    for _ in range(3):
        orc.step()
    dom.step(3)
    for name, g in dom.state().items():
        assert bit_identical_fraction(g, orc.get(name)) == 1.0, (case, tune, name)
