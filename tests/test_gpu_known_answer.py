"""End-to-end known answer on the GPU: the reference's default run
(256x256x256 on 4x6 ranks = one 256x256x252 domain, Static() initial state,
9 CFL steps) must print the published state and DT (assets/diff.iter9.png)."""
import pytest

import known_answers as ka
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from oracle import oracle as om

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lrot", [False, True], ids=["default", "lRot"])
def test_known_answer_iteration9(lrot):
    K = ka.KA_LROT if lrot else ka.KA_DEFAULT
    cfg = ka.default_config(lRot=lrot)
    desc = Descriptor(cfg)
    dom = Domain(desc)
    dom.set_state(ka.default_state(cfg, om))
    dts = [dom.step(1) for _ in range(9)]
    st = dom.state()
    for n in ("RU", "RV", "RW", "RO", "TT"):
        assert "%20.14E" % st[n][10, 10, 10] == "%20.14E" % K[n], n
    if not lrot:
        assert "%20.14E" % dts[-1] == "%20.14E" % K["DT9"]
    else:
        assert "%20.14E" % dts[-1] == "5.53534681618432E-04"
    dt10, w, _ = dom.cfl()
    dom.close()
