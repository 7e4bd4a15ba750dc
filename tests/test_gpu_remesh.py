"""The remesh round trip through the C-ABI (remeshing.cxx:1937-2035): run,
download, build a new mesh on the host and move the fields onto it, create a
context on the new mesh, upload the fields, des_remesh_finish (volume,
volume_old = volume, mass, shape functions, bottom coord0 reset), carry the
scalars over and continue -- against the oracle taking the same sequence on
the same new mesh and fields.

The host side here is a stand-in for the reference's mesher and field
transfer (out of this library's scope): a differently jittered box mesh of
the same domain, element fields from the nearest old element centroid and
node fields from the nearest old node."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import scenarios as SC
from dynearthsol_amd import model as M
from oracle import OracleState

pytestmark = pytest.mark.gpu

ELEM_FIELDS = ("stress", "strain", "plstrain", "delta_plstrain", "stressyy", "strain_rate")
NODE_FIELDS = ("vel", "temperature", "coord0")


def transfer(old_model, old, new_model):
    """Nearest-neighbour transfer of the downloaded fields onto the new mesh."""
    oc = old["coord"][old_model.conn].mean(axis=1)
    nc = new_model.coord[new_model.conn].mean(axis=1)
    ke = cKDTree(oc).query(nc)[1]
    kn = cKDTree(old["coord"]).query(new_model.coord)[1]
    out = {k: np.ascontiguousarray(old[k][ke]) for k in ELEM_FIELDS}
    out.update({k: np.ascontiguousarray(old[k][kn]) for k in NODE_FIELDS})
    return out


@pytest.mark.parametrize("name,option", [("maxwell3", 1), ("evp2", 0), ("mat2", 11)])
def test_remesh_round_trip_matches_oracle(name, option):
    from dynearthsol_amd.lib import DesModel
    m = SC.build(name)
    g = DesModel(m)
    g.run(15)
    old = g.download()
    dt, t, steps = g.scalars()
    g.close()
    # the "new mesh": the same domain, more elements, another jitter
    n = SC.build(name, nelem=5000 if m.ndims == 3 else 3600, mesh_seed=7)
    assert n.nelem != m.nelem
    f = transfer(m, old, n)
    # device: new context, fields, remesh()'s tail, scalars
    h = DesModel(n, init=False)
    h.upload(**f)
    h.remesh_finish(option)
    h._check(h.lib.des_set_scalars(h.ctx, dt, t, steps, m.compensation_pressure))
    # oracle: the same sequence
    o = OracleState(n, init=False)
    for k, v in f.items():
        o.f[k][:] = v
    o.call("compute_volume")
    o.f["volume_old"][:] = o.f["volume"]
    o.call("compute_mass")
    o.call("compute_shape_fn")
    if option in (1, 2, 11):
        bottom = (n.bcflag & M.BOUNDZ0) != 0
        o.f["coord0"][bottom, n.ndims - 1] = -n.params.zlength
    o.S.dt, o.S.time, o.S.steps, o.S.compensation_pressure = dt, t, steps, m.compensation_pressure
    got = h.download()
    for k in ("volume", "volume_old", "mass", "tmass", "volume_n", "coord0", "shpdx", "shpdz"):
        assert SC.rel_err(got[k], o.f[k]) <= 1e-13, k
    h.run(10)
    o.step(10)
    got = h.download()
    for k in got:
        if k == "stressyy" and not n.params.is_plane_strain:
            continue
        assert SC.rel_err(got[k], o.f[k]) <= 1e-10, k
    assert h.scalars()[2] == steps + 10
