"""The reference's own output files (tests/golden/frame_*.npz, written by its
Output::write and Output::write_checkpoint, tests/golden/make_frames.py) read
with this package's reader: every array of the state comes back bit for bit,
the element diagnostics equal the oracle's material functions, and the
checkpoint holds time, compensation pressure, segments, volume_old and (plane
strain) stressyy.  Pins the reader, and through it the writer's format that
the GPU tests compare byte for byte."""
import ctypes as C
import glob
import os

import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import output
from oracle import OracleState

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FRAMES = sorted(glob.glob(os.path.join(HERE, "frame_*.npz")))
NAMES = {"evp3": 500, "psevp2": 400, "mat3": 500}


def fixture_files(path, tmpdir):
    """The fixture's files written out under tmpdir: (prefix, frame, npz)."""
    z = np.load(path, allow_pickle=False)
    frame = int(z["frame"])
    prefix = os.path.join(str(tmpdir), "golden")
    for key, ext in (("file_save", f"save.{frame:06d}"), ("file_chkpt", f"chkpt.{frame:06d}"),
                     ("file_info", "info")):
        with open(f"{prefix}.{ext}", "wb") as f:
            f.write(z[key].tobytes())
    return prefix, frame, z


def scenario_of(path):
    name = os.path.basename(path)[len("frame_"):-len(".npz")]
    return name, SC.build(name, nelem=NAMES[name])


@pytest.mark.parametrize("path", FRAMES, ids=[os.path.basename(p) for p in FRAMES])
def test_reader_reads_the_reference_files(path, tmp_path):
    name, m = scenario_of(path)
    prefix, frame, z = fixture_files(path, tmp_path)
    nd, rev, fr = output.read_frame(f"{prefix}.save.{frame:06d}", m.nnode, m.nelem)
    assert nd == m.ndims and rev == 3
    for key, field in [("coordinate", "coord"), ("velocity", "vel"), ("temperature", "temperature"),
                       ("plastic strain", "plstrain"), ("plastic strain-rate", "delta_plstrain"),
                       ("strain-rate", "strain_rate"), ("strain", "strain"), ("stress", "stress"),
                       ("force", "force"), ("coord0", "coord0")]:
        assert np.array_equal(fr[key], z["state_" + field]), key
    assert np.array_equal(fr["connectivity"], m.conn)
    assert np.array_equal(fr["bcflag"].view(np.uint32), m.bcflag)
    # element diagnostics: the oracle's rho and visc of the same state, exactly
    o = OracleState(m, init=False)
    for k in o.f:
        o.f[k][:] = z["state_" + k]
    rho_fn = getattr(o.lib, f"o{m.ndims}_mat_rho")
    visc_fn = getattr(o.lib, f"o{m.ndims}_mat_visc")
    rho = np.array([rho_fn(C.byref(o.S), C.byref(o.params), int(e)) for e in range(m.nelem)])
    visc = np.array([visc_fn(C.byref(o.S), C.byref(o.params), int(e)) for e in range(m.nelem)])
    assert np.array_equal(fr["density"], rho)
    assert np.array_equal(fr["viscosity"], visc)
    q, w = o.worst_quality()
    assert np.min(fr["mesh quality"]) == q and int(np.argmin(fr["mesh quality"])) == w
    mat = np.zeros(m.nelem) if m.elemmarkers is None else np.argmax(m.elemmarkers, axis=1).astype(float)
    assert np.array_equal(fr["material"], mat)
    # the checkpoint
    ck = output.read_checkpoint(f"{prefix}.chkpt.{frame:06d}")
    dt, time_, cp, steps = z["scalars"]
    assert ck["time compensation_pressure"][0] == time_ and ck["time compensation_pressure"][1] == cp
    assert np.array_equal(ck["segment"], z["segment"]) and np.array_equal(ck["segflag"], z["segflag"])
    assert np.array_equal(ck["volume_old"], z["state_volume_old"])
    assert ("stressyy" in ck) == bool(m.params.is_plane_strain)
    if m.params.is_plane_strain:
        assert np.array_equal(ck["stressyy"], z["state_stressyy"])
    # the .info line
    (row,) = output.read_info(f"{prefix}.info")
    assert row[0] == frame and row[1] == int(steps) and row[5:] == (m.nnode, m.nelem, len(z["segflag"]))
    assert abs(row[2] - time_) <= 1e-6 * time_ and abs(row[3] - dt) <= 1e-4 * dt


def test_frame_fixtures_present():
    assert len(FRAMES) >= 3
