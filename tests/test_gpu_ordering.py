"""The device numbering of the mesh (des_set_mesh_order: Hilbert-curve order of
positions, or the host's order) is invisible at the interface: both give
bitwise equal fields, scalars, quality indices, output frames, batched
models and gradients, because every transfer permutes rows back and every
node sums its elements in ascending host order."""
import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi
from dynearthsol_amd import model as M

pytestmark = pytest.mark.gpu


def _make(m, order, **kw):
    from dynearthsol_amd.lib import DesModel, set_mesh_order
    set_mesh_order(order)
    try:
        return DesModel(m, **kw)
    finally:
        set_mesh_order(1)


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", ["evp3", "evp2", "mat3", "mat2", "bc3", "bc2", "damp23", "fields3",
                                  "elastic3", "psevp2", "dyn3", "dyn2", "dynmat3", "dynmat2", "viscous3"])
def test_fused_run_is_independent_of_numbering(name):
    m = SC.build(name)
    print(name, m.nelem, "elements", flush=True)
    g0, g1 = _make(m, 0), _make(m, 1)
    _same(g0.download(), g1.download())
    g0.run(25)
    g1.run(25)
    _same(g0.download(), g1.download())
    assert g0.scalars() == g1.scalars()


def _shuffled(m, seed):
    """The same model with its elements renumbered at random."""
    import copy
    m2 = copy.deepcopy(m)
    perm = np.random.default_rng(seed).permutation(m.nelem)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.nelem)
    m2.conn = np.ascontiguousarray(m.conn[perm])
    for k in ("stress0", "strain0", "stressyy0", "plstrain0", "elem_visc", "elem_cohesion", "elemmarkers"):
        v = getattr(m, k)
        if v is not None:
            setattr(m2, k, np.ascontiguousarray(v[perm]))
    b, e, f = m.bfacets
    m2.bfacets = (b, inv[e].astype(e.dtype), f)
    return m2


@pytest.mark.parametrize("name", ["evp3", "evp2"])
def test_direct_gathers_equal_staged_gathers(name):
    """Passes A and E stage a block's node records in LDS when the block's
    256 elements touch at most 256 distinct nodes, and gather directly
    otherwise.  Randomly numbered elements in the host's order overflow every
    block (direct gathers); the curve order of the same mesh stages them."""
    m = _shuffled(SC.build(name), 3)
    blocks = [len(np.unique(m.conn[b:b + 256])) for b in range(0, m.nelem - 255, 256)]
    print(name, m.nelem, "elements; distinct nodes per block in host order >=", min(blocks), flush=True)
    assert min(blocks) > 256
    g0, g1 = _make(m, 0), _make(m, 1)
    g0.run(25)
    g1.run(25)
    _same(g0.download(), g1.download())
    g0.run(10, abi.DES_MIXED)
    g1.run(10, abi.DES_MIXED)
    _same(g0.download(), g1.download())


@pytest.mark.parametrize("name", ["evp3", "mat2"])
def test_per_function_steps_are_independent_of_numbering(name):
    m = SC.build(name)
    g0, g1 = _make(m, 0), _make(m, 1)
    g0.step(6)
    g1.step(6)
    _same(g0.download(), g1.download())


def test_mixed_precision_is_independent_of_numbering():
    m = SC.build("evp3")
    g0, g1 = _make(m, 0), _make(m, 1)
    g0.run(20, abi.DES_MIXED)
    g1.run(20, abi.DES_MIXED)
    _same(g0.download(), g1.download())


def test_quality_ties_resolve_to_the_first_host_index():
    """An unjittered box has many elements of equal quality: the index
    returned is the first in host order, as the reference's loop keeps it."""
    m = SC.build("evp3", jitter=0.0)
    g0, g1 = _make(m, 0), _make(m, 1)
    q0, w0 = g0.worst_quality()
    q1, w1 = g1.worst_quality()
    assert (q0, w0) == (q1, w1)
    # the first strict minimum of the host-order loop
    from oracle import OracleState
    qo, wo = OracleState(m).worst_quality()
    assert wo == w1 and abs(q1 - qo) <= 1e-14 * qo


def test_tiny_elements_report_the_first_host_index():
    m = SC.build("evp3", nelem=3000)
    nd = m.ndims
    smallest = m.params.smallest_size * 0.118 * m.params.resolution ** nd
    for order in (0, 1):
        g = _make(m, order)
        vol = g.download(["volume"])["volume"]
        vol[[1700, 900, 2400]] = smallest * 0.5
        g.upload(volume=vol)
        assert g.bad_mesh_quality() == (3, 900)


def test_output_frames_are_identical(tmp_path):
    m = SC.build("mat3")
    for order in (0, 1):
        g = _make(m, order)
        g.run(10)
        g.write_frame(tmp_path / f"o{order}", 0, run_time=1.0)
    a = (tmp_path / "o0.save.000000").read_bytes()
    b = (tmp_path / "o1.save.000000").read_bytes()
    assert a == b


def test_gradient_is_independent_of_numbering():
    m = M.config_adjoint("plastic", nelem_target=800, log_visc=(19.0, 21.0))
    SC._perturb_stress(m, 30e6, 11)
    m.elem_cohesion[:] *= 0.2
    top = (m.bcflag & 32) != 0
    cb = np.zeros((m.nnode, 3))
    cb[top, 2] = np.random.default_rng(0).standard_normal(int(top.sum()))
    out = []
    for order in (0, 1):
        g = _make(m, order)
        g.adj_forward(12, 5)
        out.append(g.adj_backward(cb))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("name", ["evp3", "psevp2", "mat3", "bc3", "dynmat3"])
@pytest.mark.parametrize("mixed", [False, True])
def test_fused_run_is_independent_of_sweep_order(name, mixed):
    """The node passes' block order (des_set_sweep_order) changes which
    workgroup computes a node and when, never how: bitwise equal runs."""
    from dynearthsol_amd.lib import DesModel, set_sweep_order
    if mixed and name == "psevp2":
        pytest.skip("mixed precision does not cover plane strain")
    m = SC.build(name)
    prec = abi.DES_MIXED if mixed else abi.DES_FP64
    print(name, m.nelem, "elements, mixed" if mixed else "elements, fp64", flush=True)
    out = []
    for order in (0, 1):
        set_sweep_order(order)
        try:
            g = DesModel(m)
            g.run(20, prec)
            out.append((g.download(), g.scalars()))
            g.close()
        finally:
            set_sweep_order(1)
    _same(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("name", ["evp3", "psevp2", "bc3", "bc2", "mat3", "dynmat3", "slant3"])
def test_fused_run_is_independent_of_force_layout(name):
    """Node-major slots addressed by list start + one-byte position (0), the
    same slots by their full 4-byte positions (2) or element-major records (1)
    for the force contributions (des_set_force_layout): pass F sums each
    node's contributions in its CSR order every way, so the runs are bitwise
    equal."""
    from dynearthsol_amd.lib import DesModel, set_force_layout
    m = SC.build(name)
    print(name, m.nelem, "elements", flush=True)
    out = []
    for layout in (0, 1, 2):
        set_force_layout(layout)
        try:
            g = DesModel(m)
            g.run(23)
            out.append((g.download(), g.scalars()))
            g.close()
        finally:
            set_force_layout(0)
    for k in (1, 2):
        _same(out[0][0], out[k][0])
        assert out[0][1] == out[k][1]


@pytest.mark.parametrize("name", ["evp3", "maxwell2", "mat3"])
def test_mixed_run_is_independent_of_slot_addressing(name):
    """The mixed schedule's FP32 slots, found from list start + one-byte
    position (0) or by full positions (2)."""
    from dynearthsol_amd.lib import DesModel, set_force_layout
    m = SC.build(name)
    print(name, m.nelem, "elements, mixed", flush=True)
    out = []
    for layout in (0, 2):
        set_force_layout(layout)
        try:
            g = DesModel(m)
            g.run(23, abi.DES_MIXED)
            out.append((g.download(), g.scalars()))
            g.close()
        finally:
            set_force_layout(0)
    _same(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_batch_is_independent_of_force_layout():
    from dynearthsol_amd.lib import DesBatch, set_force_layout
    models = [SC.build("evp3"), SC.build("bc2"), SC.build("maxwell3", nelem=1500)]
    out = []
    for layout in (0, 1, 2):
        set_force_layout(layout)
        try:
            bt = DesBatch(models)
            bt.run(21)
            bt.sync()
            out.append([bt.download(k) for k in range(len(models))])
            bt.close()
        finally:
            set_force_layout(0)
    for k in range(len(models)):
        _same(out[0][k], out[1][k])
        _same(out[0][k], out[2][k])
