"""The remesh trigger bad_mesh_quality (remeshing.cxx:1882-1934): oracle on
crafted meshes, and the HIP reduction against the oracle."""
import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import model as M
from oracle import OracleState


def _state(name="evp3", ne=800):
    m = SC.build(name, nelem=ne)
    return m, OracleState(m)


@pytest.mark.parametrize("name", ["evp3", "evp2"])
def test_good_mesh_passes(name):
    m, o = _state(name)
    assert o.bad_mesh_quality() == (0, 0)


@pytest.mark.parametrize("name", ["evp3", "evp2"])
def test_tiny_element_reason_3(name):
    m, o = _state(name)
    nd = m.ndims
    sizefactor = 0.118 if nd == 3 else 0.433
    smallest = m.params.smallest_size * sizefactor * m.params.resolution ** nd
    o.f["volume"][[17, 9, 40]] = smallest * 0.5
    assert o.bad_mesh_quality() == (3, 9)


def test_bottom_drift_reason_2_only_with_bottom_options():
    m, o = _state("evp3")
    bottom = np.nonzero(m.bcflag & M.BOUNDZ0)[0]
    n = bottom[len(bottom) // 2]
    o.f["coord"][n, 2] -= 0.3 * m.params.resolution       # > 0.25 resolution
    assert o.bad_mesh_quality(remeshing_option=0) == (0, 0)
    for opt in (1, 2, 11):
        assert o.bad_mesh_quality(remeshing_option=opt) == (2, n)
    assert o.bad_mesh_quality(remeshing_option=1, max_boundary_distortion=0.5) == (0, 0)


def test_distorted_element_reason_1():
    m, o = _state("evp3")
    q, w = o.worst_quality()
    m.params.min_quality = q ** (1 / 3) * 1.01
    assert o.bad_mesh_quality() == (1, w)
    m.params.min_quality = q ** (1 / 3) * 0.99
    assert o.bad_mesh_quality() == (0, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["evp3", "evp2"])
def test_gpu_bad_mesh_quality_matches_oracle(name):
    from dynearthsol_amd.lib import DesModel
    m, o = _state(name, ne=3000)
    g = DesModel(m)
    assert g.bad_mesh_quality() == o.bad_mesh_quality() == (0, 0)
    q, w = o.worst_quality()
    cube = q ** (1 / 3) if m.ndims == 3 else q
    m.params.min_quality = cube * 1.01
    g2 = DesModel(m)
    assert g2.bad_mesh_quality() == o.bad_mesh_quality() == (1, w)
    m.params.min_quality = 0.4
    # a bottom node pushed down, uploaded to both
    bottom = np.nonzero(m.bcflag & M.BOUNDZ0)[0]
    n = bottom[len(bottom) // 3]
    o.f["coord"][n, m.ndims - 1] -= 0.4 * m.params.resolution
    g3 = DesModel(m)
    g3.upload(coord=o.f["coord"])
    assert g3.bad_mesh_quality(remeshing_option=1) == o.bad_mesh_quality(remeshing_option=1) == (2, n)
