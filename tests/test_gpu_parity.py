"""HIP library against the oracle, through the C-ABI.

Bar: every field within 1e-10 normwise relative error of the oracle (FP64;
the only sources of difference are the device libm's pow/exp/sin/tan/atan2/cos
against glibc's).  Integer and index data never leave the host.
"""
import numpy as np
import pytest

import scenarios as SC
from oracle import OracleState

pytestmark = pytest.mark.gpu

TOL = 1e-10

FUNCS = ["compute_volume", "compute_mass", "compute_shape_fn", "update_temperature",
         "update_strain_rate", "compute_dvoldt", "compute_edvoldt", "update_stress",
         "nmd_stress", "update_force", "update_velocity", "apply_vbcs", "update_mesh",
         "rotate_stress"]


def _gpu(model, init=True):
    from dynearthsol_amd.lib import DesModel
    return DesModel(model, init=init)


def _compare(g, o, names=None, tol=TOL):
    got = g.download()
    bad = {}
    for k in (names or got):
        if k == "stressyy" and not o.params.is_plane_strain:
            continue
        e = SC.rel_err(got[k], o.f[k])
        if e > tol:
            bad[k] = e
    assert not bad, bad
    return got


def _upload_oracle_state(g, o):
    g.upload(elem_visc=o.elem_visc, elem_cohesion=o.elem_cohesion, **{k: v for k, v in o.f.items()})
    g.lib.des_set_scalars(g.ctx, o.S.dt, o.S.time, o.S.steps, o.S.compensation_pressure)


@pytest.mark.parametrize("name", ["evp3", "evp2", "bc3", "mat2"])
def test_init_matches_oracle(name):
    m = SC.build(name)
    o = OracleState(m)
    g = _gpu(m)
    _compare(g, o)
    dt, t, st = g.scalars()
    assert abs(dt - o.S.dt) <= 1e-14 * o.S.dt and st == 0


@pytest.mark.parametrize("name", ["evp3", "psevp2", "bc3", "bc2", "mat3", "fields3", "dyn3", "maxwell2"])
@pytest.mark.parametrize("fn", FUNCS)
def test_each_function(name, fn):
    m = SC.build(name, nelem=1500)
    o = OracleState(m)
    o.step(3)
    g = _gpu(m, init=False)
    _upload_oracle_state(g, o)
    g.call(fn)
    o.call(fn)
    _compare(g, o)


ALL_STEP = ["evp3", "evp2", "ep3", "ps2", "psevp2", "maxwell3", "maxwell2", "elastic3", "elastic2",
            "bc3", "bc2", "damp23", "damp22", "mat3", "mat2", "fields3", "dyn3"]
# 3D models whose elements yield: the Cardano eigen solver (dsyevc3.c) cancels
# catastrophically for near-isotropic stress, so 1-ulp differences between the
# device libm and glibc in atan2/cos/sin grow over steps
PLASTIC_3D = {"evp3", "ep3", "bc3", "damp23", "mat3", "fields3", "dyn3"}


@pytest.mark.parametrize("name", ALL_STEP)
@pytest.mark.parametrize("path", ["step", "run"])
def test_one_step_from_same_state(name, path):
    """One whole step (per-function kernels or fused) from the state after 3
    oracle steps: every field within 1e-10."""
    m = SC.build(name)
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


def _ref_spread(m, nsteps):
    """How far the reference's stock build (3x3-C with -ffast-math) drifts from
    its strict build over nsteps: max relative difference over the fields."""
    from oracle import ref_available
    if not (ref_available(3) and ref_available(3, strict=True)):
        return None
    a = OracleState(m)
    b = a.copy()
    a.ref_run(nsteps, strict=True)
    b.ref_run(nsteps)
    fields = max(SC.rel_err(b.f[k], a.f[k]) for k in ("stress", "vel", "coord", "plstrain", "force"))
    return max(fields, abs(b.S.dt - a.S.dt) / abs(a.S.dt))


def _check_trajectory(g, o, name, nsteps):
    """Fields after nsteps against the oracle; returns the relative bound that
    the time step size is held to."""
    if name not in PLASTIC_3D:
        _compare(g, o)
        return 1e-9
    got = g.download()
    err = max(SC.rel_err(got[k], o.f[k]) for k in ("stress", "vel", "coord", "plstrain", "force"))
    spread = _ref_spread(o.model, nsteps)
    bound = 1e-5 if spread is None else max(10 * spread, 1e-9)
    print(f"{name}: {nsteps} steps, GPU vs oracle {err:.3e}, reference build spread {spread}")
    assert err < bound, (err, spread)
    return bound


@pytest.mark.parametrize("name", ALL_STEP)
def test_steps_match_oracle(name):
    """des_step (per-function kernels) over 12 steps, across a dt update.
    1e-10 for every model without 3D yielding; 3D plastic models stay within
    ten times the spread between the reference's own two builds."""
    m = SC.build(name)
    o = OracleState(m)
    g = _gpu(m)
    for _ in range(12):
        g.step(1)
    o.step(12)
    bound = _check_trajectory(g, o, name, 12)
    dt, t, st = g.scalars()
    assert st == 12 and abs(dt - o.S.dt) <= bound * abs(o.S.dt)


@pytest.mark.parametrize("name", ["evp3", "evp2", "ps2", "maxwell3", "elastic2", "bc3", "bc2",
                                  "damp23", "mat3", "mat2", "fields3"])
def test_fused_run_matches_oracle(name):
    """des_run (fused six-pass schedule) in two chunks, 7 + 6 steps."""
    m = SC.build(name)
    o = OracleState(m)
    g = _gpu(m)
    g.run(7)
    g.run(6)
    o.step(13)
    _check_trajectory(g, o, name, 13)
    dt, t, st = g.scalars()
    assert st == 13


def test_fused_equals_per_function_bitwise():
    """Fused and per-function paths perform the same arithmetic in the same
    order: their results are bitwise equal."""
    m = SC.build("evp3")
    g1 = _gpu(m)
    g2 = _gpu(m)
    for _ in range(15):
        g1.step(1)
    g2.run(15)
    a, b = g1.download(), g2.download()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_run_is_deterministic():
    m = SC.build("mat3")
    g1, g2 = _gpu(m), _gpu(m)
    g1.run(20)
    g2.run(20)
    a, b = g1.download(), g2.download()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_worst_quality_matches_oracle():
    m = SC.build("evp3")
    o = OracleState(m)
    g = _gpu(m)
    g.run(5)
    o.step(5)
    q, w = g.worst_quality()
    qo, wo = o.worst_quality()
    assert w == wo and abs(q - qo) <= 1e-14 * qo
