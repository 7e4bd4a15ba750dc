"""The oracle's averaged fields (Output::average_fields, output.cxx:211-251,
and the averaged arrays of Output::write) pinned by states whose averages
are known in closed form."""
import numpy as np

import scenarios as SC
from dynearthsol_amd import model as M
from oracle import AveragedFields, OracleState


def _static_model():
    """Elastic, no gravity, every wall fixed, uniform initial stress: nothing
    moves, and the stress stays what it was."""
    cfg = dict(SC.BASE)
    cfg.update({"mat.rheology_type": "elastic", "control.gravity": 0.0,
                "bc.vbc_x0": 1, "bc.vbc_val_x0": 0.0, "bc.vbc_x1": 1, "bc.vbc_val_x1": 0.0,
                "bc.vbc_z0": 1, "bc.vbc_val_z0": 0.0, "bc.vbc_z1": 1, "bc.vbc_val_z1": 0.0,
                "mesh.xlength": 20e3, "mesh.zlength": 10e3, "mesh.resolution": 2e3})
    m = M.Model(cfg, 2, nelem_target=400)
    m.stress0[:] = np.array([-3e6, -5e6, 1e6])
    return m


def test_static_stress_averages_to_interval_over_interval_plus_one():
    """The interval's first step stores the stress and the other I - 1 steps
    add theirs: I samples, divided by I + 1 (output.cxx:136-141)."""
    m = _static_model()
    o = OracleState(m)
    I = 5
    av = AveragedFields(I)
    for _ in range(2 * I):
        o.step(1)
        av.after_step(o)
    assert np.array_equal(o.f["stress"], np.broadcast_to(m.stress0, o.f["stress"].shape))
    fr = av.frame(o)
    expect = (5 * m.stress0) * (1.0 / 6)
    assert np.allclose(fr["stress averaged"], expect, rtol=1e-15, atol=0)
    assert np.all(fr["velocity averaged"] == 0) and np.all(fr["strain-rate"] == 0)
    # dt of the .info line: (time - time0) / I, time0 taken at the interval's first step
    assert np.isclose(fr["dt"], o.S.dt * (I - 1) / I, rtol=1e-14)


def test_walls_moving_at_fixed_speed_average_to_that_speed():
    """x walls pulled at +-V: their displacement over the interval is V times
    the elapsed time, so the averaged velocity is V there."""
    m = SC.build("evp2", nelem=600)
    o = OracleState(m)
    av = AveragedFields(4)
    for _ in range(8):
        o.step(1)
        av.after_step(o)
    va = av.frame(o)["velocity averaged"]
    x1 = (m.bcflag & M.BOUNDX1) != 0
    x0 = (m.bcflag & M.BOUNDX0) != 0
    V = m.params.vbc_values[1]
    assert np.allclose(va[x1, 0], V, rtol=1e-9, atol=0)
    assert np.allclose(va[x0, 0], -V, rtol=1e-9, atol=0)
