"""Averaged fields on the device (des_set_averaging / des_average_fields /
des_write_frame_averaged) against the oracle's Output::average_fields, and
the fused driver against the per-function one."""
import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi, output
from oracle import AveragedFields, OracleState

pytestmark = pytest.mark.gpu

AVG_KEYS = ("velocity averaged", "plastic strain-rate", "strain-rate", "stress averaged")


@pytest.mark.parametrize("name", ["maxwell3", "evp2", "elastic3", "mat3"])
def test_averaged_frame_matches_oracle(name, tmp_path):
    from dynearthsol_amd.lib import DesModel
    m = SC.build(name)
    I, n = 5, 12
    g = DesModel(m)
    g.set_averaging(I)
    g.run(n)
    g.write_frame_averaged(tmp_path / "avg", 1, run_time=2.0)
    o = OracleState(m)
    av = AveragedFields(I)
    for _ in range(n):
        o.step(1)
        av.after_step(o)
    ref = av.frame(o)
    _, _, fr = output.read_frame(tmp_path / "avg.save.000001", m.nnode, m.nelem)
    for k in AVG_KEYS:
        err = SC.rel_err(fr[k], ref[k])
        print(name, k, f"{err:.2e}", flush=True)
        assert err < 1e-8, k
    info = output.read_info(tmp_path / "avg.info")
    assert abs(info[-1][3] - ref["dt"]) <= 1e-4 * ref["dt"]      # written with %12.4e
    # the instantaneous fields of the same frame are the device state
    got = g.download(["stress", "strain"])
    assert np.array_equal(fr["stress"], got["stress"]) and np.array_equal(fr["strain"], got["strain"])


def test_fused_and_per_function_averages_are_identical(tmp_path):
    """The fused driver applies the rotation it still owes on the fly; the
    per-function driver has applied it: the sums are bitwise the same."""
    from dynearthsol_amd.lib import DesModel
    m = SC.build("evp3")
    g1, g2 = DesModel(m), DesModel(m)
    g1.set_averaging(4)
    g2.set_averaging(4)
    for _ in range(10):
        g1.step(1)
        g1.average_fields()
    g2.run(10)
    g1.write_frame_averaged(tmp_path / "a", 0)
    g2.write_frame_averaged(tmp_path / "b", 0)
    assert (tmp_path / "a.save.000000").read_bytes() == (tmp_path / "b.save.000000").read_bytes()


def test_averaging_errors():
    from dynearthsol_amd.lib import DesError, DesModel
    g = DesModel(SC.build("evp3"))
    with pytest.raises(DesError):
        g.write_frame_averaged("/tmp/never", 0)
    with pytest.raises(DesError):
        g.average_fields()
    g.set_averaging(5)
    with pytest.raises(DesError):
        g.run(3, abi.DES_MIXED)
