"""des_create's input checks return the reference's status codes (1 for bad
parameters, 10 for an inconsistent mesh; README:139-144) before any device
call, and leave no context behind."""
import copy
import ctypes as C
import os

import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi, lib

pytestmark = pytest.mark.skipif(not os.path.exists(lib.LIB_PATH), reason="library not built")


def _create(m, mutate_mesh=None, mutate_params=None):
    h = lib.load()
    p = copy.deepcopy(m.params)
    ms = m.mesh_struct()
    if mutate_mesh:
        mutate_mesh(ms)
    if mutate_params:
        mutate_params(p)
    ctx = C.c_void_p()
    rc = h.des_create(C.byref(p), C.byref(ms), C.byref(ctx))
    assert not ctx.value          # nothing allocated on a rejected input
    return rc


@pytest.fixture(scope="module")
def model():
    return SC.build("evp3")


def test_node_index_out_of_range(model):
    conn = np.array(model.conn, dtype=np.int32)
    conn[5, 2] = model.nnode

    def bad(ms):
        ms.connectivity = abi.iptr(conn)
    assert _create(model, bad) == abi.DES_ERR_MESH


def test_negative_node_index(model):
    conn = np.array(model.conn, dtype=np.int32)
    conn[0, 0] = -1

    def bad(ms):
        ms.connectivity = abi.iptr(conn)
    assert _create(model, bad) == abi.DES_ERR_MESH


def test_empty_mesh(model):
    def bad(ms):
        ms.nelem = 0
    assert _create(model, bad) == abi.DES_ERR_MESH


def test_too_many_elements_for_32_bit_indices(model):
    """nelem * (ndims + 1) >= 2^31 is refused before the connectivity is read."""
    def bad(ms):
        ms.nelem = (1 << 31) // 4
    assert _create(model, bad) == abi.DES_ERR_MESH


def test_boundary_facet_of_a_missing_element(model):
    b, e, f = model.bfacets
    e2 = np.array(e, dtype=np.int32)
    e2[0] = model.nelem

    def bad(ms):
        ms.bfacet_elem = abi.iptr(e2)
    assert _create(model, bad) == abi.DES_ERR_MESH


def test_bad_dimension(model):
    def bad(p):
        p.ndims = 4
    assert _create(model, mutate_params=bad) == abi.DES_ERR_INPUT


def test_materials_without_markers(model):
    def bad(p):
        p.nmat = 2
    assert _create(model, mutate_params=bad) == abi.DES_ERR_INPUT
