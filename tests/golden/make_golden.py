"""Write the golden fixtures: small models stepped by the reference's own step
functions (oracle/_ref/libref{2d,3d_strict}.so, compiled from the reference
sources by oracle/Makefile), saved as inputs and expected outputs.

    python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import scenarios as SC  # noqa: E402
from oracle import OracleState  # noqa: E402

# (scenario, elements, steps recorded, reference steps taken before the
# recorded input state)
CASES = [("evp3", 500, 20, 0), ("evp2", 400, 20, 0), ("mat3", 500, 15, 0), ("bc2", 400, 20, 0),
         ("psevp2", 400, 20, 0), ("maxwell3", 500, 15, 0), ("viscous3", 500, 15, 0),
         ("viscous2", 400, 20, 0), ("dynmat3", 500, 1, 9), ("bc3", 500, 1, 9)]


def main():
    for name, ne, nsteps, pre in CASES:
        m = SC.build(name, nelem=ne)
        o = OracleState(m)
        if pre:
            o.ref_run(pre, strict=True)
        out = {
            "nsteps": np.array(nsteps),
            "params": np.frombuffer(bytes(m.params), dtype=np.uint8).copy(),
            "ndims": np.array(m.ndims),
            "conn": m.conn, "bcflag": m.bcflag,
            "bf_bdry": m.bfacets[0], "bf_elem": m.bfacets[1], "bf_face": m.bfacets[2],
            "bnormals": m.bnormals, "edge_vectors": m.edge_vectors, "edge_valid": m.edge_valid,
            "markers_per_element": np.array(m.markers_per_element),
            "scalars_in": np.array([o.S.dt, o.S.time, o.S.compensation_pressure, o.S.steps]),
        }
        if m.elemmarkers is not None:
            out["elemmarkers"] = m.elemmarkers
        for k, v in o.f.items():
            out["in_" + k] = v.copy()
        o.ref_run(nsteps, strict=True)
        for k, v in o.f.items():
            out["out_" + k] = v.copy()
        out["scalars_out"] = np.array([o.S.dt, o.S.time, o.S.compensation_pressure, o.S.steps])
        path = os.path.join(HERE, f"{name}_from{pre}.npz" if pre else f"{name}.npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
