"""Reader of the reference's binary output frames (binaryio.cxx:16-36): a
4096-byte ASCII header -- "# DynEarthSol ndims=N revision=R", then
"name<TAB>offset" lines -- followed by the raw arrays.  Field shapes follow
the reference's own reader (Dynearthsol.py:74-100)."""
import numpy as np

HEADERLEN = 4096

NSTR_FIELDS = {"strain", "strain-rate", "stress", "stress averaged"}
ELEM_FIELDS = {"density", "material", "mesh quality", "plastic strain", "plastic strain-rate",
               "viscosity", "edvoldt", "volume"}
NODE_VEC_FIELDS = {"coordinate", "velocity", "velocity averaged", "force", "coord0"}
NODE_FIELDS = {"bcflag", "temperature", "mass", "tmass", "volume_n"}


def read_header(path):
    with open(path, "rb") as f:
        header = f.read(HEADERLEN)
    lines = header.split(b"\n")
    first = lines[0].split()
    assert first[0] == b"#" and first[1] == b"DynEarthSol", first
    ndims = int(first[2].split(b"=")[1])
    revision = int(first[3].split(b"=")[1])
    pos = {}
    for line in lines[1:]:
        if not line or line[0] == 0:
            break
        name, p = line.split(b"\t")
        pos[name.decode("ascii")] = int(p)
    return ndims, revision, pos


def read_frame(path, nnode, nelem):
    """All arrays of one .save file, shaped as the reference's reader does."""
    ndims, revision, pos = read_header(path)
    nstr = 3 if ndims == 2 else 6
    out = {}
    with open(path, "rb") as f:
        for name, p in pos.items():
            dtype = np.int32 if name in ("connectivity", "bcflag") else np.float64
            if name in NSTR_FIELDS:
                shape = (nelem, nstr)
            elif name in ELEM_FIELDS:
                shape = (nelem,)
            elif name == "connectivity":
                shape = (nelem, ndims + 1)
            elif name in NODE_VEC_FIELDS:
                shape = (nnode, ndims)
            elif name in NODE_FIELDS:
                shape = (nnode,)
            else:
                raise KeyError(name)
            f.seek(p)
            out[name] = np.fromfile(f, dtype=dtype, count=int(np.prod(shape))).reshape(shape)
    return ndims, revision, out


def read_info(path):
    """Rows of the .info table: frame, steps, time, dt, run_time, nnode, nelem, nseg."""
    rows = []
    with open(path) as f:
        for line in f:
            v = line.split()
            rows.append((int(v[0]), int(v[1]), float(v[2]), float(v[3]), float(v[4]),
                         int(v[5]), int(v[6]), int(v[7])))
    return rows


CHKPT_DTYPES = {"time compensation_pressure": np.float64, "segment": np.int32, "segflag": np.int32,
                "volume_old": np.float64, "stressyy": np.float64}


def read_arrays(path):
    """Every array of a file in the binary format as raw bytes, by name (an
    array runs from its offset to the next one's, the last to the end)."""
    _, _, pos = read_header(path)
    with open(path, "rb") as f:
        data = f.read()
    names = sorted(pos, key=pos.get)
    ends = [pos[n] for n in names[1:]] + [len(data)]
    return {n: data[pos[n]:end] for n, end in zip(names, ends)}


def read_checkpoint(path):
    """The arrays of a .chkpt file (Output::write_checkpoint,
    output.cxx:254-276): the known ones typed, the others (marker sets) as
    raw bytes."""
    ndims, _, _ = read_header(path)
    out = {}
    for name, raw in read_arrays(path).items():
        dt = CHKPT_DTYPES.get(name)
        a = np.frombuffer(raw, dtype=dt).copy() if dt is not None else raw
        if name == "segment":
            a = a.reshape(-1, ndims)
        out[name] = a
    return out
