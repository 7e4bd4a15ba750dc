"""Run-time failures come back as the reference's status codes instead of
std::exit: dt <= 0 from an inverted element is 11 (geometry.cxx:304-308),
an unknown rheology 1 (rheology.cxx:636-639).  des_run is asynchronous: a
failure inside the step is reported by the next des_sync (or by the next
compute_dt, every 10 steps)."""
import copy

import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi

pytestmark = pytest.mark.gpu


def test_inverted_element_gives_status_11():
    from dynearthsol_amd.lib import DesError, DesModel
    m = copy.deepcopy(SC.build("evp3"))
    conn = np.array(m.conn, dtype=np.int32)
    conn[7, [0, 1]] = conn[7, [1, 0]]          # negative volume
    m.conn = conn
    with pytest.raises(DesError) as ex:
        g = DesModel(m)
        g.run(10)
    print("status", ex.value.code, ex.value, flush=True)
    assert ex.value.code == abi.DES_ERR_RUNTIME


@pytest.mark.parametrize("precision", [abi.DES_FP64, abi.DES_MIXED])
def test_unknown_rheology_gives_status_1(precision):
    from dynearthsol_amd.lib import DesError, DesModel
    m = copy.deepcopy(SC.build("evp3"))
    m.params.rheol_type = 6                     # no such rheology
    with pytest.raises(DesError) as ex:
        g = DesModel(m)
        g.run(3, precision)
        g.scalars()                             # des_get_scalars reports it, as des_sync does
    print("status", ex.value.code, ex.value, flush=True)
    assert ex.value.code == abi.DES_ERR_INPUT


def test_context_still_usable_after_a_refused_adjoint():
    """A refused call leaves the model as it was: the run continues."""
    from dynearthsol_amd.lib import DesError, DesModel
    m = SC.build("evp3")
    g = DesModel(m)
    g.run(5)
    before = g.download(["coord"])["coord"]
    with pytest.raises(DesError):
        g.adj_forward(2, 1)                    # evp3 has no per-element viscosity
    assert np.array_equal(g.download(["coord"])["coord"], before)
    g.run(5)
    assert g.scalars()[2] == 10


def test_batch_reports_the_failing_model_only():
    from dynearthsol_amd.lib import DesBatch, DesError
    good, bad = SC.build("evp3"), copy.deepcopy(SC.build("elastic3"))
    bad.params.rheol_type = 6                   # no such rheology
    bt = DesBatch([good, bad])
    bt.run(3)
    assert bt.scalars(0)[2] == 3
    with pytest.raises(DesError) as ex:
        bt.scalars(1)
    assert ex.value.code == abi.DES_ERR_INPUT
