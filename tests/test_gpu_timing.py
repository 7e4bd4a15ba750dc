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
    for v in range(6):
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
    assert 0 <= v < 6 and all(x > 0 for x in ms) and ms[v] == min(ms)
    with pytest.raises(DesError):
        g.set_tuning(6)


VARIANTS = {"pass_a": 3, "pass_c": 6, "pass_d": 5, "pass_e": 3, "pass_f": 5, "pass_b": 3}
MIXED_VARIANTS = dict(VARIANTS, pass_e=2)


@pytest.mark.parametrize("name,prec", [("evp3", abi.DES_FP64), ("evp3", abi.DES_MIXED), ("mat2", abi.DES_FP64),
                                       ("bc3", abi.DES_FP64)])
def test_every_variant_of_every_pass_gives_identical_results(name, prec):
    """Each tunable pass run in each of its variants (the others at their
    defaults) leaves bitwise the same state after 20 steps."""
    from dynearthsol_amd.lib import DesModel, TUNE_PASSES
    m = SC.build(name)
    ref = DesModel(m)
    ref.set_tuning_all(prec, [0, 1 if prec == abi.DES_FP64 else 0, 0, 0, 0, 0])
    ref.run(20, prec)
    a = ref.download()
    for p in TUNE_PASSES:
        nv = (MIXED_VARIANTS if prec == abi.DES_MIXED else VARIANTS)[p]
        for v in range(nv):
            choice = [0, 1 if prec == abi.DES_FP64 else 0, 0, 0, 0, 0]
            choice[TUNE_PASSES.index(p)] = v
            g = DesModel(m)
            g.set_tuning_all(prec, choice)
            g.run(20, prec)
            b = g.download()
            for k in a:
                assert np.array_equal(a[k], b[k]), (p, v, k)
            assert g.scalars() == ref.scalars()
            g.close()


@pytest.mark.parametrize("prec", [abi.DES_FP64, abi.DES_MIXED])
def test_autotuner_tunes_every_pass(prec):
    """A run of 16+ steps tunes every pass: the choice is the variant with
    the smallest measured mean, and the tuned run equals an untuned one."""
    from dynearthsol_amd.lib import DesModel
    m = SC.build("evp3")
    g, h = DesModel(m), DesModel(m)
    assert all(v == -1 for v, _ in g.tuning_all(prec).values())
    g.run(20, prec)
    h.set_tuning_all(prec, [0, 1 if prec == abi.DES_FP64 else 0, 0, 0, 0, 0])
    h.run(20, prec)
    t = g.tuning_all(prec)
    print(prec, {p: (v, [round(x, 4) for x in ms]) for p, (v, ms) in t.items()}, flush=True)
    for p, (v, ms) in t.items():
        nv = (MIXED_VARIANTS if prec == abi.DES_MIXED else VARIANTS)[p]
        assert 0 <= v < nv, (p, v)
        assert all(x > 0 for x in ms[:nv]) and ms[v] == min(ms[:nv]), (p, ms)
    a, b = g.download(), h.download()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    with pytest.raises(Exception):
        g.set_tuning_all(prec, [3, 0, 0, 0, 0, 0])


def test_batched_groups_tune_every_pass():
    from dynearthsol_amd.lib import DesBatch
    bt = DesBatch([SC.build(n) for n in ("evp3", "maxwell3", "evp2", "viscous2")])
    bt.run(20, abi.DES_MIXED)
    bt.sync()
    for nd in (2, 3):
        t = bt.tuning_all(nd, abi.DES_MIXED)
        print(nd, {p: v for p, (v, _) in t.items()}, flush=True)
        assert all(v >= 0 for v, _ in t.values()), t
