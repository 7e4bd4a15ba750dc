"""Mixed-precision fused step (DES_MIXED) against the FP64 step.

FP32 deviatoric stress, strain, plastic strain, strain rate and rheology
arithmetic; FP64 pressure, coordinates, volumes, mass, force accumulation and
eigen-return.  The error bar is 1e-4 relative (normwise per field) on
velocity and stress against the FP64 run; other fields are reported.
"""
import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi, model as M
from oracle import OracleState

pytestmark = pytest.mark.gpu

FIELDS = ("vel", "stress", "coord", "strain", "plstrain", "temperature", "force")


def _pair(m):
    from dynearthsol_amd.lib import DesModel
    return DesModel(m), DesModel(m)


def _errors(a, b):
    return {k: SC.rel_err(a[k], b[k]) for k in FIELDS}


@pytest.mark.parametrize("name", ["evp3", "evp2", "maxwell3", "elastic3", "fields3", "mat3"])
def test_mixed_close_to_fp64(name):
    m = SC.build(name)
    g64, g32 = _pair(m)
    g64.run(300)
    g32.run(300, abi.DES_MIXED)
    e = _errors(g32.download(), g64.download())
    print(name, {k: f"{v:.2e}" for k, v in e.items()})
    assert e["vel"] < 1e-4 and e["stress"] < 1e-4, e


def test_mixed_chunks_and_state_roundtrip():
    """Mixed chunks leave an FP64 state that continues seamlessly."""
    m = SC.build("evp3")
    g1, g2 = _pair(m)
    g1.run(40, abi.DES_MIXED)
    g2.run(15, abi.DES_MIXED)
    g2.run(25, abi.DES_MIXED)
    e = _errors(g1.download(), g2.download())
    assert e["vel"] < 1e-5 and e["stress"] < 1e-5, e
    dt, t, st = g2.scalars()
    assert st == 40


def test_mixed_long_run_10k_steps():
    """The mixed-precision configuration (per-element viscosity and cohesion)
    at reduced size, 10,000 steps: velocity and stress within 1e-4 of FP64."""
    m = M.config_benchmark("3d-mixed", nelem_target=20000)
    g64, g32 = _pair(m)
    g64.run(10000)
    g32.run(10000, abi.DES_MIXED)
    a, b = g32.download(), g64.download()
    e = _errors(a, b)
    print("10k steps mixed vs fp64:", {k: f"{v:.2e}" for k, v in e.items()})
    assert e["vel"] < 1e-4 and e["stress"] < 1e-4, e


@pytest.mark.parametrize("name,nsteps", [("evp3", 300), ("maxwell3", 300), ("evp2", 300)])
def test_fp64_sensitivity_floor(name, nsteps):
    """How far two FP64 runs drift apart when the initial stress differs by
    1e-12 relative: the model's own noise floor for trajectory comparisons."""
    from dynearthsol_amd.lib import DesModel
    m = SC.build(name)
    a = DesModel(m)
    b = DesModel(m)
    s0 = m.stress0 * (1 + 1e-12 * (2 * M.splitmix_uniform(5, m.nelem)[:, None] - 1))
    b.upload(stress=s0)
    a.run(nsteps)
    b.run(nsteps)
    e = _errors(b.download(), a.download())
    print(f"FP64 floor {name} {nsteps} steps:", {k: f"{v:.2e}" for k, v in e.items()})
    assert np.isfinite(e["vel"])


def test_mixed_against_oracle_short():
    m = SC.build("evp3")
    o = OracleState(m)
    from dynearthsol_amd.lib import DesModel
    g = DesModel(m)
    g.run(50, abi.DES_MIXED)
    o.step(50)
    got = g.download()
    e = {k: SC.rel_err(got[k], o.f[k]) for k in FIELDS}
    print("mixed vs oracle, 50 steps:", {k: f"{v:.2e}" for k, v in e.items()})
    assert e["vel"] < 1e-4 and e["stress"] < 1e-4, e
