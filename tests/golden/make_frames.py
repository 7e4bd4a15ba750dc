"""Write the golden output fixtures: small models stepped by the reference's
own step functions, then written by the reference's own Output::write and
Output::write_checkpoint (oracle/_ref/libref{2,3}d_io.so, compiled from the
reference sources by oracle/Makefile ref-io).  Each fixture holds the state
written (inputs) and the bytes of the .save, .chkpt and .info files
(expected outputs).

    python tests/golden/make_frames.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import scenarios as SC  # noqa: E402
from dynearthsol_amd import model as M  # noqa: E402
from oracle import OracleState  # noqa: E402

CASES = [("evp3", 500, 12, 3), ("psevp2", 400, 12, 2), ("mat3", 500, 7, 5)]   # scenario, elements, steps, frame


def segments(m):
    """The boundary facets as segments (node indices) with their flag bit."""
    nof = M.NODE_OF_FACET[m.ndims]
    b, e, f = m.bfacets
    seg = np.array([[m.conn[ee, nof[ff, j]] for j in range(m.ndims)] for ee, ff in zip(e, f)], dtype=np.int32)
    return seg.reshape(-1, m.ndims), (1 << b).astype(np.int32)


def main():
    for name, ne, nsteps, frame in CASES:
        m = SC.build(name, nelem=ne)
        o = OracleState(m)
        o.ref_run(nsteps, strict=True)
        seg, flg = segments(m)
        out = {"frame": np.array(frame), "segment": seg, "segflag": flg,
               "scalars": np.array([o.S.dt, o.S.time, o.S.compensation_pressure, o.S.steps])}
        for k, v in o.f.items():
            out["state_" + k] = v.copy()
        with tempfile.TemporaryDirectory() as d:
            base = os.path.join(d, "golden")
            o.ref_write_output(base, frame, seg, flg)
            for ext in (f"save.{frame:06d}", f"chkpt.{frame:06d}", "info"):
                with open(f"{base}.{ext}", "rb") as fh:
                    out["file_" + ext.split(".")[0]] = np.frombuffer(fh.read(), dtype=np.uint8).copy()
        path = os.path.join(HERE, f"frame_{name}.npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
