"""Output frames in the reference's binary format, written from device state
(des_write_frame) and read back with the format's own layout."""
import ctypes as C

import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import output
from oracle import OracleState

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["evp3", "mat2"])
def test_frame_roundtrip(name, tmp_path):
    from dynearthsol_amd.lib import DesModel
    m = SC.build(name)
    g = DesModel(m)
    prefix = tmp_path / "model"
    g.write_frame(prefix, 0)
    g.run(12)
    g.write_frame(prefix, 1, run_time=1.5)
    nd, rev, fr = output.read_frame(f"{prefix}.save.000001", m.nnode, m.nelem)
    assert nd == m.ndims and rev == 3
    got = g.download()
    for key, field in [("coordinate", "coord"), ("velocity", "vel"), ("temperature", "temperature"),
                       ("plastic strain", "plstrain"), ("plastic strain-rate", "delta_plstrain"),
                       ("strain-rate", "strain_rate"), ("strain", "strain"), ("stress", "stress"),
                       ("force", "force"), ("coord0", "coord0")]:
        assert np.array_equal(fr[key], got[field]), key
    assert np.array_equal(fr["connectivity"], m.conn)
    assert np.array_equal(fr["bcflag"].view(np.uint32), m.bcflag)
    # element diagnostics against the oracle's material functions at the same state
    o = OracleState(m, init=False)
    for k in o.f:
        if k in got:
            o.f[k][:] = got[k]
    rho_fn = getattr(o.lib, f"o{m.ndims}_mat_rho")
    visc_fn = getattr(o.lib, f"o{m.ndims}_mat_visc")
    idx = np.arange(0, m.nelem, 97)
    rho = np.array([rho_fn(C.byref(o.S), C.byref(o.params), int(e)) for e in idx])
    visc = np.array([visc_fn(C.byref(o.S), C.byref(o.params), int(e)) for e in idx])
    assert np.max(np.abs(fr["density"][idx] - rho) / rho) < 1e-13
    assert np.max(np.abs(fr["viscosity"][idx] - visc) / visc) < 1e-12
    assert np.all((fr["mesh quality"] > 0) & (fr["mesh quality"] <= 1.0 + 1e-12))
    if m.elemmarkers is not None:
        assert np.array_equal(fr["material"], np.argmax(m.elemmarkers, axis=1).astype(float))
    info = output.read_info(f"{prefix}.info")
    assert [r[0] for r in info] == [0, 1] and info[1][1] == 12
    assert info[1][5] == m.nnode and info[1][6] == m.nelem
