"""Batched models (des_batch_*): every model of a batch must end bitwise equal
to the same model run on its own context -- the batch launches the same
kernels over concatenated meshes -- and so inherit the single-model parity
with the oracle.  One model is also checked against the oracle directly."""
import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi
from dynearthsol_amd import model as M
from oracle import OracleState

pytestmark = pytest.mark.gpu

# rheologies, boundary conditions (facet tractions in bc2/bc3), plane strain,
# dynamic elastic dt, and a model without elasticity (no rotation)
MIX = ["evp3", "maxwell3", "bc3", "damp23", "elastic3", "evp2", "bc2", "ps2", "viscous2",
       "maxwell2", "damp22", "elastic2"]


def _single(m, steps, precision=abi.DES_FP64):
    from dynearthsol_amd.lib import DesModel
    g = DesModel(m)
    for n in steps:
        g.run(n, precision)
    return g.download(), g.scalars()


def _assert_same(a, b, tag):
    for k in a:
        assert np.array_equal(a[k], b[k]), (tag, k, float(np.max(np.abs(a[k] - b[k]))))


def _check_batch(models, steps, precision=abi.DES_FP64):
    from dynearthsol_amd.lib import DesBatch
    bt = DesBatch(models)
    for n in steps:
        bt.run(n, precision)
    bt.sync()
    for k, m in enumerate(models):
        ref, sc = _single(m, steps, precision)
        _assert_same(bt.download(k), ref, k)
        assert bt.scalars(k) == sc, k
    return bt


def test_batch_equals_single_models_fp64():
    """12 different models, 7 + 6 steps across a dt update."""
    models = [SC.build(n) for n in MIX]
    print("batch of", [(n, m.nelem) for n, m in zip(MIX, models)], flush=True)
    _check_batch(models, [7, 6])


def test_batch_equals_single_models_mixed():
    models = [SC.build(n) for n in ["evp3", "maxwell3", "elastic3", "evp2", "maxwell2", "viscous2"]]
    _check_batch(models, [12], abi.DES_MIXED)


def test_batch_per_element_fields_and_materials():
    """Per-element viscosity / cohesion (all models of the batch carry them)
    and two-material models with marker counts."""
    _check_batch([SC.build("fields3"), SC.build("fields3", nelem=2500)], [11])
    _check_batch([SC.build("mat3"), SC.build("mat2"), SC.build("mat2", nelem=1800)], [11])


def test_batch_rejects_partial_per_element_fields():
    from dynearthsol_amd.lib import DesBatch, DesError
    bt = DesBatch([SC.build("fields3"), SC.build("evp3")])
    with pytest.raises(DesError):
        bt.run(1)


def test_batch_model_against_oracle():
    """Model 1 of a batch, one step from the initial state, against the oracle
    within 1e-10 (the single-step bar of the single-model path)."""
    from dynearthsol_amd.lib import DesBatch
    models = [SC.build("evp3"), SC.build("evp2"), SC.build("maxwell3", nelem=2000)]
    bt = DesBatch(models)
    o = OracleState(models[1])
    bt.run(1)
    o.step(1)
    got = bt.download(1)
    for k in ("coord", "vel", "force", "stress", "strain", "plstrain", "temperature", "volume", "mass"):
        assert SC.rel_err(got[k], o.f[k]) < 1e-10, k


def test_batch_ensemble_dt_and_quality():
    """A small uneven ensemble: per-model dt and worst quality equal the
    single-model values."""
    from dynearthsol_amd.lib import DesBatch, DesModel
    models = M.config_ensemble(n2d=5, n3d=5, size_scale=0.05)
    bt = DesBatch(models)
    dt0 = bt.compute_dt()
    bt.run(10)
    dt1 = bt.compute_dt()
    q, w = bt.worst_quality()
    for k, m in enumerate(models):
        g = DesModel(m)
        assert g.compute_dt() == dt0[k], k
        g.run(10)
        assert g.compute_dt() == dt1[k], k
        assert g.worst_quality() == (q[k], w[k]), k
