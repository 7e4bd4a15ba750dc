"""The persistent launch (des_set_launch_mode / des_batch_set_launch_mode 1):
a whole run of fused steps in one cooperative kernel with grid barriers
between the passes.  It runs the per-pass kernels' own block code in the same
order per element and node, so every field and scalar must come out bitwise
equal to the per-pass launches -- and so inherit their parity with the
oracle -- over runs of odd and even length (volume buffers swapped inside the
kernel), across compute_dt updates, for single models and for a batch whose
2D and 3D models share the one launch."""
import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi

pytestmark = pytest.mark.gpu

# rheologies, plasticity with the deferred return list, facet tractions,
# plane strain, dynamic dt, two materials, per-element fields, slant walls
SINGLE = ["evp3", "ep3", "psevp2", "bc3", "bc2", "mat3", "mat2", "fields3", "dynmat3",
          "viscous2", "maxwell3", "slant3", "slanttb2"]


def _run(m, mode, runs, precision):
    from dynearthsol_amd.lib import DesModel
    g = DesModel(m)
    g.set_launch_mode(mode)
    for n in runs:
        g.run(n, precision)
    out = g.download(), g.scalars()
    g.close()
    return out


def _same(a, b, tag):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), (tag, k, float(np.max(np.abs(a[k] - b[k]))))


@pytest.mark.parametrize("name", SINGLE)
def test_persistent_equals_per_pass_fp64(name):
    m = SC.build(name)
    print(name, m.nelem, "elements", flush=True)
    runs = [7, 12, 1]          # odd and even lengths; steps 10 and 20 update dt
    a, sa = _run(m, 0, runs, abi.DES_FP64)
    b, sb = _run(m, 1, runs, abi.DES_FP64)
    _same(a, b, name)
    assert sa == sb


@pytest.mark.parametrize("name", ["evp3", "mat3", "maxwell2", "fields3"])
def test_persistent_equals_per_pass_mixed(name):
    m = SC.build(name)
    print(name, m.nelem, "elements, mixed", flush=True)
    a, sa = _run(m, 0, [13, 8], abi.DES_MIXED)
    b, sb = _run(m, 1, [13, 8], abi.DES_MIXED)
    _same(a, b, name)
    assert sa == sb


@pytest.mark.parametrize("precision", [abi.DES_FP64, abi.DES_MIXED])
def test_persistent_batch_equals_per_pass(precision):
    """2D and 3D models of uneven size in one launch."""
    from dynearthsol_amd.lib import DesBatch
    names = ["evp3", "maxwell3", "elastic3", "evp2", "maxwell2", "viscous2"]
    models = [SC.build(n) for n in names]
    models.append(SC.build("evp3", nelem=900))
    print("batch of", [m.nelem for m in models], flush=True)
    out = []
    for mode in (0, 1):
        bt = DesBatch(models)
        bt.set_launch_mode(mode)
        bt.run(9, precision)
        bt.run(6, precision)
        bt.sync()
        out.append([(bt.download(k), bt.scalars(k)) for k in range(len(models))])
        bt.close()
    for k in range(len(models)):
        _same(out[0][k][0], out[1][k][0], k)
        assert out[0][k][1] == out[1][k][1], k


def test_persistent_batch_of_one_dimension():
    from dynearthsol_amd.lib import DesBatch
    models = [SC.build("bc2"), SC.build("evp2", nelem=700)]
    out = []
    for mode in (0, 1):
        bt = DesBatch(models)
        bt.set_launch_mode(mode)
        bt.run(11)
        bt.sync()
        out.append([(bt.download(k), bt.scalars(k)) for k in range(len(models))])
        bt.close()
    for k in range(len(models)):
        _same(out[0][k][0], out[1][k][0], k)
        assert out[0][k][1] == out[1][k][1], k


def test_launch_mode_rejects_unknown_modes():
    from dynearthsol_amd.lib import DesModel, DesError
    g = DesModel(SC.build("evp2"))
    with pytest.raises(DesError):
        g.set_launch_mode(2)
    g.close()
