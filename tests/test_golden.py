"""The oracle against the committed golden fixtures (tests/golden/*.npz):
small models stepped by the reference's own compiled step functions
(tests/golden/make_golden.py).  Bit-exact."""
import glob
import os
import types

import numpy as np
import pytest

from dynearthsol_amd import abi
from oracle import OracleState

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# step trajectories (frame_*.npz hold output files; tests/test_golden_frames.py)
FIXTURES = sorted(p for p in glob.glob(os.path.join(HERE, "*.npz"))
                  if not os.path.basename(p).startswith("frame_"))


def load_fixture(path):
    z = np.load(path, allow_pickle=False)
    m = types.SimpleNamespace()
    m.params = abi.DesParams.from_buffer_copy(z["params"].tobytes())
    nd = int(z["ndims"])
    m.ndims, m.npe, m.nstr = nd, nd + 1, 6 if nd == 3 else 3
    m.conn = z["conn"].astype(np.int32)
    m.coord = z["in_coord"]
    m.nnode, m.nelem = len(m.coord), len(m.conn)
    m.bcflag = z["bcflag"].astype(np.uint32)
    m.bfacets = (z["bf_bdry"].astype(np.int32), z["bf_elem"].astype(np.int32),
                 z["bf_face"].astype(np.int32))
    m.bnormals, m.edge_vectors = z["bnormals"], z["edge_vectors"]
    m.edge_valid = z["edge_valid"].astype(np.int32)
    m.markers_per_element = int(z["markers_per_element"])
    m.elemmarkers = z["elemmarkers"].astype(np.int32) if "elemmarkers" in z else None
    m.elem_visc = m.elem_cohesion = None
    return m, z


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_golden(path):
    m, z = load_fixture(path)
    o = OracleState(m, init=False)
    for k in o.f:
        o.f[k][:] = z["in_" + k]
    dt, t, cp, st = z["scalars_in"]
    o.S.dt, o.S.time, o.S.compensation_pressure, o.S.steps = dt, t, cp, int(st)
    assert o.step(int(z["nsteps"])) == 0
    for k in o.f:
        if k == "stressyy" and not m.params.is_plane_strain:
            continue
        assert np.array_equal(o.f[k], z["out_" + k]), k
    assert o.S.dt == z["scalars_out"][0] and o.S.steps == int(z["scalars_out"][3])


def test_fixtures_present():
    assert len(FIXTURES) >= 4
