"""Parity at the workloads' own sizes (GPU).

* The GPU MHD drop-in configuration (MHD + rotation, 256x256x128, the
  benchmark's state): after 10 and after 100 CFL steps bit-identical to the
  oracle (the stated bar is 1e-12 per field), which runs on the host's cores
  (about a third of a second a step).
* The benchmark's own workload (MHD + rotation, 512x512x256, Ra 1e6, the
  device-built state): after 3 CFL steps the FP64 form bit-identical to the
  oracle (about 3 s a step on 16 host cores), the mixed form within 1e-5
  per field of it.
* The mixed-precision workload (MHD + rotation, 512x512x256, Ra 1e6): after
  100 CFL steps from the same device-built state, every field of the mixed
  form within 1e-5 of the FP64 form (max|g-c|/max|c| over the interior incl.
  walls), total mass within 1e-12 and total energy within 1e-9 relative.
"""
import numpy as np
import pytest

from gpu_common import bit_identical_fraction, relerr
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.initial import stratified_config
from oracle import oracle as om

pytestmark = pytest.mark.gpu


def test_headline_size_10_and_100_steps():
    print("fullsize: FP64 MHD+rot 256x256x128, 10 and 100 CFL steps vs oracle", flush=True)
    desc = Descriptor(stratified_config(256, 256, 128, lMag=True, lRot=True))
    dom = Domain(desc)
    dom.init_stratified(seed=12345)
    st = dom.state()
    orc = om.Oracle(desc, threads=16)
    orc.set_state(st)
    orc.communicate()
    del st
    done = 0
    for upto in (10, 100):
        dts = [orc.step() for _ in range(upto - done)]
        assert dom.step(upto - done) == dts[-1], upto
        done = upto
        for name, g in dom.state().items():
            c = orc.get(name)
            e = relerr(g, c)
            assert e <= 1e-12, (upto, name, e)
            assert bit_identical_fraction(g, c) == 1.0, (upto, name, e)
        print(f"  {upto} steps: bit-identical", flush=True)
    orc.close()
    dom.close()


def test_benchmark_size_3_steps_against_the_oracle():
    print("fullsize: FP64 and mixed MHD+rot 512x512x256 Ra 1e6, 3 CFL steps vs oracle", flush=True)
    size = (512, 512, 256)
    cfg = stratified_config(*size, Ra=1e6, lMag=True, lRot=True)
    desc = Descriptor(cfg)
    dom = Domain(desc)
    dom.init_stratified(seed=4242)
    orc = om.Oracle(desc, threads=16)
    orc.set_state(dom.state())
    orc.communicate()
    dts = [orc.step() for _ in range(3)]
    print("  oracle done", flush=True)
    assert dom.step(3) == dts[-1]
    for name, g in dom.state().items():
        c = orc.get(name)
        e = relerr(g, c)
        assert e <= 1e-12, (name, e)
        assert bit_identical_fraction(g, c) == 1.0, (name, e)
    dom.close()
    print("  FP64: bit-identical", flush=True)
    mx = Domain(Descriptor(stratified_config(*size, Ra=1e6, lMag=True, lRot=True, precision=1)))
    mx.init_stratified(seed=4242)
    mx.step(3)
    for name, g in mx.state().items():
        e = relerr(g, orc.get(name))
        print(f"  mixed {name}: {e:.3e}", flush=True)
        assert e <= 1e-5, (name, e)
    mx.close()
    orc.close()


def test_mixed_full_size_100_steps():
    print("fullsize: mixed vs FP64, MHD+rot 512x512x256, 100 CFL steps", flush=True)
    size = (512, 512, 256)
    doms = []
    for prec in (0, 1):
        cfg = stratified_config(*size, Ra=1e6, lMag=True, lRot=True, precision=prec)
        dom = Domain(Descriptor(cfg))
        dom.init_stratified(seed=4242)
        dom.step(100)
        doms.append(dom)
    g64, gmx = doms[0].diagnostics(), doms[1].diagnostics()
    s64 = doms[0].state()
    doms[0].close()
    smx = doms[1].state()
    doms[1].close()
    for name in s64:
        e = relerr(smx[name], s64[name])
        print(f"  {name}: {e:.3e}", flush=True)
        assert e <= 1e-5, (name, e)
    assert abs(gmx["mass"] - g64["mass"]) <= 1e-12 * abs(g64["mass"])
    assert abs(gmx["energy"] - g64["energy"]) <= 1e-9 * abs(g64["energy"])
    assert np.isfinite(gmx["max_divB"])
