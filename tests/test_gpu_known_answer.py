"""End-to-end known answer on the GPU: the reference's default run
(256x256x256 on 4x6 ranks = one 256x256x252 domain, Static() initial state,
9 CFL steps) must print the published state and DT (assets/diff.iter9.png).
The initial state comes from the product's Static() (mhd3d_static_state), so
the published values are reproduced by product code alone; the oracle's
Static() is checked against it in tests/test_static_state.py."""
import pytest

import known_answers as ka
from mhd3d.desc import Descriptor
from mhd3d.domain import Domain
from mhd3d.initial import reference_static_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lrot", [False, True], ids=["default", "lRot"])
def test_known_answer_iteration9(lrot):
    K = ka.KA_LROT if lrot else ka.KA_DEFAULT
    cfg = ka.default_config(lRot=lrot)
    desc = Descriptor(cfg)
    dom = Domain(desc)
    dom.set_state(reference_static_state(cfg, ka.NP, ka.NPE_Y, ka.NPE_Z))
    dts = [dom.step(1) for _ in range(9)]
    st = dom.state()
    for n in ("RU", "RV", "RW", "RO", "TT"):
        assert "%20.14E" % st[n][10, 10, 10] == "%20.14E" % K[n], n
    if not lrot:
        assert "%20.14E" % dts[-1] == "%20.14E" % K["DT9"]
    else:
        assert "%20.14E" % dts[-1] == "5.53534681618432E-04"
    dom.close()
