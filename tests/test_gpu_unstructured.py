"""The step on unstructured meshes: Delaunay tetrahedralizations (and
triangulations) of jittered point lattices, whose node degrees vary as a
mesher's do (3D: 6 to 36 elements per node, against the structured Kuhn
mesh's few fixed degrees) -- the gathers' worst case for locality.  Same bar
as the structured scenarios: one fused step from the oracle's state within
1e-10, free runs of non-amplifying models over 12 steps within 1e-10, fused
equal to per-function and both device numberings equal, bitwise."""
import numpy as np
import pytest

import scenarios as SC
from oracle import OracleState
from test_gpu_parity import _compare, _upload_oracle_state

pytestmark = pytest.mark.gpu


def _gpu(m, init=True, order=1):
    from dynearthsol_amd.lib import DesModel, set_mesh_order
    set_mesh_order(order)
    try:
        return DesModel(m, init=init)
    finally:
        set_mesh_order(1)


@pytest.mark.parametrize("name", ["evp3", "maxwell3", "bc3", "evp2", "mat2", "slant3"])
@pytest.mark.parametrize("path", ["step", "run"])
def test_one_step_on_delaunay_mesh(name, path):
    m = SC.build(name, mesh="delaunay")
    o = OracleState(m)
    o.step(3)
    g = _gpu(m, init=False)
    _upload_oracle_state(g, o)
    if path == "step":
        g.step(1)
    else:
        g.run(1)
    o.step(1)
    _compare(g, o)


@pytest.mark.parametrize("name", ["maxwell3", "evp2", "viscous3", "slant2"])
def test_free_run_on_delaunay_mesh(name):
    m = SC.build(name, mesh="delaunay")
    o = OracleState(m)
    g = _gpu(m)
    g.run(7)
    g.run(5)
    o.step(12)
    _compare(g, o)
    dt, t, st = g.scalars()
    assert st == 12 and abs(dt - o.S.dt) <= 1e-12 * o.S.dt


def test_fused_per_function_and_numberings_agree_on_delaunay_mesh():
    m = SC.build("evp3", mesh="delaunay")
    g1, g2, g0 = _gpu(m), _gpu(m), _gpu(m, order=0)
    for _ in range(10):
        g1.step(1)
    g2.run(10)
    g0.run(10)
    a, b, c = g1.download(), g2.download(), g0.download()
    for k in a:
        assert np.array_equal(a[k], b[k]) and np.array_equal(b[k], c[k]), k
