"""HIP library against the oracle, through the C-ABI.

Bar: every field within 1e-10 normwise relative error of the oracle (FP64;
the only sources of difference are the device libm's pow/exp/sin/tan/atan2/cos
against glibc's), dt within 1e-12.  Integer and index data never leave the
host.  Free-running trajectories of the models that amplify ulp-level
differences (3D yielding, undamped dynamic plasticity) are held to a fixed
1e-8 over a measured per-model number of steps, and every step of their
12-step trajectories is checked at 1e-10 from the oracle's state (re-seeded
chain).
"""
import numpy as np
import pytest

import scenarios as SC
from oracle import OracleState

pytestmark = pytest.mark.gpu

TOL = 1e-10
DT_TOL = 1e-12

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


@pytest.mark.parametrize("name", ["evp3", "evp2", "bc3", "mat2", "dyn3", "dyn2", "dynmat3", "dynmat2"])
def test_init_matches_oracle(name):
    """init() and its compute_dt.  The dynamic models (dyn*, dynmat*) take
    the elastic time step from bulkm and rho of the reference's element 0
    (geometry.cxx:297), which the curve numbering puts elsewhere on the
    device; with a depth-dependent density (alpha != 0) or two materials,
    element 0's rho differs from that of the device's first element."""
    m = SC.build(name)
    o = OracleState(m)
    g = _gpu(m)
    _compare(g, o)
    dt, t, st = g.scalars()
    assert abs(dt - o.S.dt) <= 1e-14 * o.S.dt and st == 0


@pytest.mark.parametrize("name", ["evp3", "mat2", "dyn3", "dyn2", "dynmat3", "dynmat2"])
@pytest.mark.parametrize("path", ["step", "run"])
def test_dt_update_matches_oracle(name, path):
    """The step that ends with compute_dt (steps % 10 == 0,
    dynearthsol.cxx:356-366), from the oracle's state after 9 steps: dt within
    1e-12, every field within 1e-10."""
    m = SC.build(name)
    o = OracleState(m)
    o.step(9)
    g = _gpu(m, init=False)
    _upload_oracle_state(g, o)
    if path == "step":
        g.step(1)
    else:
        g.run(1)
    o.step(1)
    _compare(g, o)
    dt, t, st = g.scalars()
    assert st == 10
    assert abs(dt - o.S.dt) <= DT_TOL * o.S.dt, (dt, o.S.dt)


@pytest.mark.parametrize("name", ["evp3", "psevp2", "bc3", "bc2", "mat3", "fields3", "dyn3", "maxwell2",
                                  "viscous2", "viscous3", "dynmat3", "slant3", "slanttb2"])
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
            "viscous3", "viscous2", "bc3", "bc2", "damp23", "damp22", "mat3", "mat2", "fields3", "dyn3",
            "dyn2", "dynmat3", "dynmat2", "slant3", "slant2", "slanttb3", "slanttb2"]
# Models whose trajectories amplify ulp-level differences (the device libm's
# pow/exp/sin/tan/atan2/cos against glibc's): 3D models whose elements yield
# (the Cardano eigen solver, dsyevc3.c, cancels catastrophically for
# near-isotropic stress), and the undamped dynamic plastic models, where
# elements sitting on the yield surface flip between returning and not (the
# oracle itself goes from a 1e-14 stress perturbation to O(1) differences in
# two steps, tests/test_oracle_math.py::test_dynamic_plastic_sensitivity).
AMPLIFYING = {"evp3", "ep3", "bc3", "damp23", "mat3", "fields3", "dyn3", "dynmat3", "dyn2", "dynmat2"}


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


# Free-running trajectories of those models are held to the fixed TRAJ_TOL
# over FREE_STEPS steps: runs over which the measured GPU-vs-oracle error
# stayed at or below a third of it on the MI355X (evp3 3.3e-9 at 4 steps,
# ep3 3.0e-9 at 6, bc3 2.1e-9 at 2, damp23 5.9e-9 at 3, mat3 at most 2.2e-9
# over 12, fields3 2.9e-9 at 4; the reference's own two builds, 3x3-C with
# and without -ffast-math, drift apart by as much: 6.4e-9 after 2 steps).
# The undamped dynamic plastic models run one free step (7e-14 measured).
# The re-seeded chain checks every step of the 12 at 1e-10.
FREE_STEPS = {"evp3": 4, "ep3": 6, "bc3": 2, "damp23": 3, "mat3": 12, "fields3": 4, "dyn3": 1,
              "dynmat3": 1, "dyn2": 1, "dynmat2": 1}
TRAJ_TOL = 1e-8
TRAJ_FIELDS = ("stress", "vel", "coord", "plstrain", "force", "temperature", "strain")


def _steps_for(name, default):
    return FREE_STEPS[name] if name in AMPLIFYING else default


def _check_trajectory(g, o, name, nsteps):
    """Fields after nsteps against the oracle: 1e-10, or the fixed TRAJ_TOL
    for the AMPLIFYING models."""
    if name not in AMPLIFYING:
        _compare(g, o)
        return
    got = g.download()
    err = {k: SC.rel_err(got[k], o.f[k]) for k in TRAJ_FIELDS}
    print(f"{name}: {nsteps} steps, GPU vs oracle " + " ".join(f"{k} {v:.2e}" for k, v in err.items()))
    assert max(err.values()) < TRAJ_TOL, err


@pytest.mark.parametrize("name", ALL_STEP)
def test_steps_match_oracle(name):
    """des_step (per-function kernels), free-running from init: 12 steps
    across a dt update at 1e-10, FREE_STEPS at 1e-8 for the AMPLIFYING
    models; dt within 1e-12."""
    m = SC.build(name)
    o = OracleState(m)
    g = _gpu(m)
    n = _steps_for(name, 12)
    for _ in range(n):
        g.step(1)
    o.step(n)
    _check_trajectory(g, o, name, n)
    dt, t, st = g.scalars()
    assert st == n and abs(dt - o.S.dt) <= DT_TOL * abs(o.S.dt)


@pytest.mark.parametrize("name", ["evp3", "evp2", "ps2", "maxwell3", "elastic2", "bc3", "bc2",
                                  "damp23", "mat3", "mat2", "fields3", "dyn2", "dynmat2", "viscous2",
                                  "viscous3", "slant3", "slant2", "slanttb3", "slanttb2"])
def test_fused_run_matches_oracle(name):
    """des_run (fused six-pass schedule) in two chunks: 13 steps at 1e-10,
    FREE_STEPS at 1e-8 for the AMPLIFYING models."""
    m = SC.build(name)
    o = OracleState(m)
    g = _gpu(m)
    n = _steps_for(name, 13)
    first = (n + 1) // 2
    g.run(first)
    if n > first:
        g.run(n - first)
    o.step(n)
    _check_trajectory(g, o, name, n)
    dt, t, st = g.scalars()
    assert st == n and abs(dt - o.S.dt) <= DT_TOL * abs(o.S.dt)


@pytest.mark.parametrize("name", sorted(AMPLIFYING))
@pytest.mark.parametrize("path", ["step", "run"])
def test_reseeded_chain_matches_oracle(name, path):
    """Every step of a 12-step trajectory of the AMPLIFYING models (one of
    them ends with compute_dt), each from the oracle's state at its start:
    every field within 1e-10 and dt within 1e-12, step after step."""
    m = SC.build(name)
    o = OracleState(m)
    g = _gpu(m, init=False)
    worst = 0.0
    for k in range(12):
        _upload_oracle_state(g, o)
        if path == "step":
            g.step(1)
        else:
            g.run(1)
        o.step(1)
        got = _compare(g, o)
        worst = max(worst, max(SC.rel_err(got[f], o.f[f]) for f in TRAJ_FIELDS))
        dt, t, st = g.scalars()
        assert st == k + 1 and abs(dt - o.S.dt) <= DT_TOL * abs(o.S.dt), (k, dt, o.S.dt)
    print(f"{name} ({path}): worst single-step error over 12 steps {worst:.2e}")


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


@pytest.mark.parametrize("name", ["slant3", "slant2", "slanttb3", "slanttb2"])
def test_slant_boundaries_take_the_slant_branch(name):
    """The slant walls' velocity BCs (bc.cxx:350-436) on the device: nodes on
    n0 alone get the prescribed normal velocity (type 1: v.n = value, type 3:
    v = value * n), intersections with the y walls and the bottom are
    projected on the edge vectors, and the whole velocity field equals the
    oracle's apply_vbcs bitwise-close (1e-10)."""
    m = SC.build(name)
    o = OracleState(m)
    o.step(2)
    g = _gpu(m, init=False)
    _upload_oracle_state(g, o)
    g.call("apply_vbcs")
    o.call("apply_vbcs")
    v = _compare(g, o, names=["vel"])["vel"]
    only_n0 = m.bcflag == (1 << 6)
    assert only_n0.sum() > 0
    vn = v[only_n0] @ m.bnormals[6]
    assert np.allclose(vn, SC.V1, rtol=1e-14, atol=0), (vn.min(), vn.max())
    if name.startswith("slanttb"):
        assert np.allclose(v[only_n0], SC.V1 * m.bnormals[6], rtol=1e-14, atol=0)
    edge = (m.bcflag & (1 << 6)) != 0
    edge &= (m.bcflag & 4) != 0                 # n0 meets y0 (both type 1): v along the edge vector
    edge &= (m.bcflag & (16 | 32 | 8)) == 0
    if name == "slant3":
        ev = m.edge_vectors[2, 6]
        ve = v[edge] @ ev
        assert edge.sum() > 0 and np.allclose(v[edge], ve[:, None] * ev[None, :], rtol=0, atol=1e-24)
