"""The measurement entry points: des_run_timed (per-pass kernel times from
dispatch-timestamped events) and the pass-C autotuner (des_get_tuning /
des_set_tuning).  Neither the timing nor the variant changes a result."""
import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi

pytestmark = pytest.mark.gpu

PASSES = ("pass_a", "pass_b", "pass_c", "pass_d", "pass_e", "pass_f")


@pytest.mark.parametrize("name,prec", [("evp3", abi.DES_FP64), ("evp3", abi.DES_MIXED),
                                       ("evp2", abi.DES_FP64), ("bc3", abi.DES_FP64),
                                       ("maxwell3", abi.DES_FP64)])
def test_run_timed_counts_and_results(name, prec):
    from dynearthsol_amd.lib import DesModel
    m = SC.build(name)
    g, h = DesModel(m), DesModel(m)
    t = g.run_timed(25, prec)
    h.run(25, prec)
    print(name, {k: (round(v[0], 4), v[1]) for k, v in t.items()}, flush=True)
    for k in PASSES:
        assert t[k][1] == 25 and t[k][0] > 0, k
    assert t["dt"][1] == 2                        # compute_dt after steps 10 and 20
    plastic = (m.params.rheol_type & abi.RH_PLASTIC) != 0
    assert t.get("pass_c_ret", (0, 0))[1] == (25 if m.ndims == 3 and plastic else 0)
    assert t.get("facet", (0, 0))[1] in (0, 25)
    a, b = g.download(), h.download()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert g.scalars() == h.scalars()


def test_pass_c_variants_give_identical_results():
    from dynearthsol_amd.lib import DesModel
    m = SC.build("evp3")
    out = []
    for v in range(4):
        g = DesModel(m)
        g.set_tuning(v)
        g.run(20)
        assert g.tuning()[0] == v
        out.append(g.download())
    for o in out[1:]:
        for k in o:
            assert np.array_equal(o[k], out[0][k]), k


def test_autotuner_keeps_the_fastest_variant():
    from dynearthsol_amd.lib import DesError, DesModel
    g = DesModel(SC.build("evp3"))
    assert g.tuning()[0] == -1
    g.run(20)
    v, ms = g.tuning()
    assert 0 <= v < 4 and all(x > 0 for x in ms) and ms[v] == min(ms)
    with pytest.raises(DesError):
        g.set_tuning(4)
