"""The oracle against the reference itself, compiled from its own sources into
oracle/_ref by oracle/Makefile (skipped where that build is absent).

* _strict: 3x3-C built without -ffast-math -- the oracle must agree bit for bit
  over whole trajectories, in 2D and 3D, for every rheology and boundary
  condition the scenarios exercise.
* stock flags (3x3-C with -ffast-math, as the reference's Makefile builds it):
  identical wherever the eigen solver is not used; in 3D plastic models the
  Cardano eigenvalues (dsyevc3.c:57-72) cancel catastrophically for
  near-isotropic stress, so reassociation under -ffast-math changes them and
  only a single step is compared, loosely.
"""
import numpy as np
import pytest

import scenarios as SC
from oracle import OracleState, ref_available, REF_CALL

# scenarios the reference can express (fields3 uses per-element viscosity and
# cohesion, an extension)
REF_SCENARIOS = [s for s in SC.SCENARIOS if not s.startswith("fields")]

needs_ref = pytest.mark.skipif(not (ref_available(2) and ref_available(3, strict=True)),
                               reason="oracle/_ref not built (needs /root/reference)")


def _compare(a, b, plane_strain, tol=0.0):
    bad = {}
    for k in a.f:
        if k == "stressyy" and not plane_strain:
            continue    # the reference leaves it undefined outside plane strain
        e = SC.rel_err(a.f[k], b.f[k])
        if e > tol:
            bad[k] = e
    assert not bad, bad
    assert a.S.dt == b.S.dt and a.S.time == b.S.time and a.S.steps == b.S.steps


@needs_ref
@pytest.mark.parametrize("name", REF_SCENARIOS)
def test_trajectory_bit_exact(name):
    m = SC.build(name)
    a = OracleState(m)
    b = a.copy()
    nsteps = 25
    assert a.step(nsteps) == 0
    b.ref_run(nsteps, strict=True)
    _compare(a, b, m.params.is_plane_strain)


@needs_ref
@pytest.mark.parametrize("name", ["evp3", "bc2", "mat3", "psevp2"])
@pytest.mark.parametrize("fn", sorted(REF_CALL))
def test_each_function_bit_exact(name, fn):
    m = SC.build(name, nelem=1500)
    a = OracleState(m)
    a.step(3)
    b = a.copy()
    a.call(fn)
    b.ref_call(fn, strict=True)
    _compare(a, b, m.params.is_plane_strain)


@needs_ref
@pytest.mark.parametrize("name", ["elastic3", "maxwell3", "evp2", "bc2"])
def test_stock_flags_identical_without_eigen(name):
    """Scenarios that never reach the 3D eigen solver (2D, or 3D without
    plasticity) are bit-exact against the reference built with its own flags."""
    m = SC.build(name)
    a = OracleState(m)
    b = a.copy()
    a.step(25)
    b.ref_run(25)
    _compare(a, b, m.params.is_plane_strain)


@needs_ref
@pytest.mark.parametrize("name", ["evp3", "ep3", "mat3", "bc3", "damp23", "dynmat3"])
def test_stock_flags_one_step_3d_plastic(name):
    """One step against the stock-flag reference: the only differences come
    from its -ffast-math eigen solver (measured: at most 1.3e-10 after one
    step on these models)."""
    m = SC.build(name)
    a = OracleState(m)
    b = a.copy()
    a.step(1)
    b.ref_run(1)
    for k in ("stress", "strain", "vel", "coord", "plstrain", "force", "temperature"):
        assert SC.rel_err(a.f[k], b.f[k]) < 1e-9, k
    assert a.S.dt == b.S.dt


@needs_ref
def test_worst_quality_matches_reference():
    m = SC.build("evp3", nelem=1500)
    a = OracleState(m)
    a.step(5)
    assert a.worst_quality() == a.ref_worst_quality()


@needs_ref
@pytest.mark.parametrize("name", ["evp3", "evp2", "maxwell3", "mat3"])
def test_uniform_element_fields_reproduce_reference(name):
    """The oracle's per-element viscosity and cohesion fields (an extension
    the reference cannot express) pinned where it can: with the viscosity
    clamp closed (visc_min = visc_max, as in the reference's CPU
    configuration) the power law evaluates to the clamp value everywhere, and
    a uniform field of that value must give the reference's run bit for bit;
    likewise a uniform cohesion field equal to cohesion0 of a one-material
    model.  A field partly outside [visc_min, visc_max] must equal its clamped
    copy (the clamp applies to the field too)."""
    m = SC.build(name)
    P = m.params
    P.visc_min = P.visc_max = 3e22
    a = OracleState(m)
    b = a.copy()
    c = a.copy()
    a.set_elem_params(elem_visc=np.full(m.nelem, 3e22),
                      elem_cohesion=np.full(m.nelem, P.cohesion0[0]) if P.nmat == 1 else None)
    a.step(15)
    b.ref_run(15, strict=True)
    _compare(a, b, P.is_plane_strain)
    if name == "evp3":
        P.visc_min, P.visc_max = 1e20, 1e23
        u = 10.0 ** (19.0 + 5.0 * np.random.default_rng(1).random(m.nelem))     # 1e19 .. 1e24
        d = c.copy()
        c.set_elem_params(elem_visc=u)
        d.set_elem_params(elem_visc=np.clip(u, 1e20, 1e23))
        c.step(10)
        d.step(10)
        _compare(c, d, False)
