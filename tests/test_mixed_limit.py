"""Why the mixed-precision velocity target (1e-4 against FP64 after 10k
steps) is out of reach of any scheme that stores the deviatoric stress in
FP32: the FP64 oracle itself, changed only by rounding its deviatoric stress
to FP32 after every step, drifts from the plain FP64 oracle in velocity by
~1e-1 within 1,000 steps (7e-2 after 10,000 steps, stress 4.8e-5; measured
in the build container), on the mixed-precision workload's model at reduced
size.  In this quasi-static dynamic-relaxation scheme the velocity is a small
residual of the force balance, and the 6e-8 relative rounding of the stress
each step is a force noise of that residual's order.  The device's mixed
path (tests/test_gpu_mixed.py) lands at the same level (5.2e-2 velocity,
7.9e-5 stress after 10k steps)."""
import numpy as np

import scenarios as SC
from dynearthsol_amd import model as M
from oracle import OracleState


def _round_deviator_to_fp32(stress, nd):
    p = stress[:, :nd].mean(axis=1, keepdims=True)
    d = stress.copy()
    d[:, :nd] -= p
    d = d.astype(np.float32).astype(np.float64)
    d[:, :nd] += p
    stress[:] = d


def test_fp32_deviator_storage_alone_exceeds_the_velocity_target():
    m = M.config_benchmark("3d-mixed", nelem_target=20000)
    a = OracleState(m)
    b = a.copy()
    for _ in range(1000):
        a.step(1)
        b.step(1)
        _round_deviator_to_fp32(b.f["stress"], 3)
    ev = SC.rel_err(b.f["vel"], a.f["vel"])
    es = SC.rel_err(b.f["stress"], a.f["stress"])
    print(f"FP32-rounded deviator, 1000 steps: velocity {ev:.2e}, stress {es:.2e}")
    assert ev > 1e-2          # two orders of magnitude above the 1e-4 target
    assert es < 2e-4          # the stress stays near the FP32 rounding level
