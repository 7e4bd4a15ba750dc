"""Output frames, checkpoints and restart on the device, against the
reference's own files and against the oracle.

* des_write_frame / des_write_checkpoint of a fixture state against the bytes
  the reference's Output::write / write_checkpoint wrote for the same state
  (tests/golden/frame_*.npz): headers identical, every array bitwise
  identical except the viscosity (the device's pow/exp against glibc's: within
  1e-14), the .info line identical but for the run time.
* restart (dynearthsol.cxx:104-218, dynearthsol_amd/restart.py) from files
  the device wrote, continued, against the oracle taking the same restart
  sequence from the same files and continuing: every field within 1e-10.
"""
import os

import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi, output
from dynearthsol_amd.restart import RestartedModel, read_restart_files
from oracle import OracleState
from test_golden_frames import FRAMES, fixture_files, scenario_of

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("path", FRAMES, ids=[os.path.basename(p) for p in FRAMES])
def test_writers_match_the_reference_files(path, tmp_path):
    from dynearthsol_amd.lib import DesModel
    name, m = scenario_of(path)
    os.makedirs(tmp_path / "ref", exist_ok=True)
    ref_prefix, frame, z = fixture_files(path, tmp_path / "ref")
    g = DesModel(m, init=False)
    g.upload(**{k[len("state_"):]: z[k] for k in z.files if k.startswith("state_")})
    dt, time_, cp, steps = z["scalars"]
    g._check(g.lib.des_set_scalars(g.ctx, float(dt), float(time_), int(steps), float(cp)))
    prefix = str(tmp_path / "dev")
    g.write_checkpoint(prefix, frame, z["segment"], z["segflag"])
    g.write_frame(prefix, frame)
    # .save: header and every array
    save = f".save.{frame:06d}"
    with open(ref_prefix + save, "rb") as a, open(prefix + save, "rb") as b:
        ra, rb = a.read(), b.read()
    assert ra[:output.HEADERLEN] == rb[:output.HEADERLEN]
    A, B = output.read_arrays(ref_prefix + save), output.read_arrays(prefix + save)
    assert list(A) == list(B)
    for k in A:
        if k == "viscosity":
            va, vb = np.frombuffer(A[k]), np.frombuffer(B[k])
            assert np.max(np.abs(va - vb) / va) <= 1e-14, k
        else:
            assert A[k] == B[k], k
    # .chkpt: byte for byte
    chk = f".chkpt.{frame:06d}"
    with open(ref_prefix + chk, "rb") as a, open(prefix + chk, "rb") as b:
        assert a.read() == b.read()
    # .info: all columns but the run time
    (ra_,), (rb_,) = output.read_info(ref_prefix + ".info"), output.read_info(prefix + ".info")
    with open(ref_prefix + ".info") as a, open(prefix + ".info") as b:
        la, lb = a.read().split(), b.read().split()
    assert la[:4] == lb[:4] and la[5:] == lb[5:]
    g.close()


@pytest.mark.parametrize("name", ["maxwell3", "evp2", "psevp2", "mat2"])
def test_restart_continues_like_the_oracle(name, tmp_path):
    """20 fused steps on the device, frame and checkpoint at step 20 (a
    multiple of 10, where the loop computes dt), restart into a new context,
    10 more steps; the oracle restarts from the same files the same way and
    steps 10 times."""
    from dynearthsol_amd.lib import DesModel
    from dynearthsol_amd.restart import restart
    m = SC.build(name)
    g = DesModel(m)
    g.run(20)
    prefix = str(tmp_path / "run")
    seg, flg = np.zeros((0, m.ndims), np.int32), np.zeros(0, np.int32)
    g.write_checkpoint(prefix, 1, seg, flg)
    g.write_frame(prefix, 1)
    h, extra = restart(m.params, prefix, 1, markers_per_element=m.markers_per_element,
                       elemmarkers=m.elemmarkers)
    assert extra["steps"] == 20
    o = oracle_restart(m, prefix, 1)
    # the restarted state and dt
    got = h.download()
    for k in ("coord", "vel", "temperature", "stress", "strain", "plstrain", "mass", "tmass", "volume",
              "volume_old"):
        assert SC.rel_err(got[k], o.f[k]) <= 1e-12, k
    dt, t, st = h.scalars()
    assert st == 20 and abs(dt - o.S.dt) <= 1e-12 * o.S.dt
    h.run(10)
    o.step(10)
    got = h.download()
    for k in got:
        if k == "stressyy" and not m.params.is_plane_strain:
            continue
        assert SC.rel_err(got[k], o.f[k]) <= 1e-10, k
    assert h.scalars()[2] == 30


def oracle_restart(m, prefix, frame):
    """The oracle through restart()'s sequence from the same files."""
    info, save, chk = read_restart_files(prefix, frame)
    r = RestartedModel(m.params, save["coordinate"], save["connectivity"], save["bcflag"].view(np.uint32),
                       m.markers_per_element, m.elemmarkers)
    o = OracleState(r, init=False)
    o.f["coord0"][:] = save["coord0"]
    o.call("compute_volume")
    o.f["volume_old"][:] = chk["volume_old"]
    o.call("compute_mass")
    o.call("compute_shape_fn")
    o.call("apply_vbcs")
    for k, key in (("vel", "velocity"), ("temperature", "temperature"), ("strain_rate", "strain-rate"),
                   ("strain", "strain"), ("stress", "stress"), ("plstrain", "plastic strain"), ("force", "force")):
        o.f[k][:] = save[key]
    if m.params.is_plane_strain:
        o.f["stressyy"][:] = chk["stressyy"]
    o.S.time, o.S.compensation_pressure = chk["time compensation_pressure"]
    o.S.steps = info[1]
    o.call("compute_dt")
    return o
