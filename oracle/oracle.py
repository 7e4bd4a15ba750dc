"""ctypes driver of the CPU oracle (liboracle.so) and of the reference harness
(_ref/libref{2,3}d[_omp].so).

TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the benchmark's
CPU-baseline leg use it to check or time; the product never imports it.
"""
import ctypes as C
import os

import numpy as np

from dynearthsol_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))

_P = C.POINTER
_d = _P(C.c_double)
_i = _P(C.c_int)


class OState(C.Structure):
    _fields_ = [
        ("nnode", C.c_int), ("nelem", C.c_int), ("nmat", C.c_int),
        ("markers_per_element", C.c_int),
        ("conn", _i), ("bcflag", _P(C.c_uint)), ("elemmarkers", _i),
        ("csr_off", _i), ("csr_ent", _i),
        ("nbfacet", C.c_int),
        ("bf_bdry", _i), ("bf_elem", _i), ("bf_face", _i),
        ("bnormals", _d), ("edge_vectors", _d), ("edge_valid", _i),
    ] + [(n, _d) for n in (
        "coord", "coord0", "vel", "force", "temperature", "mass", "tmass", "volume_n", "ntmp",
        "volume", "volume_old", "edvoldt", "plstrain", "delta_plstrain", "dpressure", "stressyy",
        "strain_rate", "strain", "stress", "shpdx", "shpdy", "shpdz",
        "elem_visc", "elem_cohesion", "ework")] + [
        ("dt", C.c_double), ("time", C.c_double), ("compensation_pressure", C.c_double),
        ("steps", C.c_int), ("status", C.c_int),
    ]


NODE_VEC = ("coord", "coord0", "vel", "force")
NODE_SCAL = ("temperature", "mass", "tmass", "volume_n", "ntmp")
ELEM_SCAL = ("volume", "volume_old", "edvoldt", "plstrain", "delta_plstrain", "dpressure", "stressyy")
ELEM_TENS = ("strain_rate", "strain", "stress")
ELEM_SHP = ("shpdx", "shpdy", "shpdz")
ALL_FIELDS = NODE_VEC + NODE_SCAL + ELEM_SCAL + ELEM_TENS + ELEM_SHP

_libs = {}


def _lib(name):
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        _libs[name] = C.CDLL(path)
    return _libs[name]


def oracle_lib():
    lib = _lib("liboracle.so")
    for nd in (2, 3):
        pf = f"o{nd}_"
        for f in ("compute_volume", "compute_shape_fn", "compute_dvoldt", "compute_edvoldt",
                  "nmd_stress", "update_strain_rate", "update_velocity", "update_coordinate",
                  "rotate_stress"):
            getattr(lib, pf + f).argtypes = [C.c_void_p]
        for f in ("compute_mass", "update_temperature", "update_force", "apply_vbcs",
                  "update_stress", "init_derived", "update_mesh"):
            getattr(lib, pf + f).argtypes = [C.c_void_p, C.c_void_p]
        getattr(lib, pf + "step").argtypes = [C.c_void_p, C.c_void_p]
        getattr(lib, pf + "step").restype = C.c_int
        getattr(lib, pf + "run").argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        getattr(lib, pf + "run").restype = C.c_int
        getattr(lib, pf + "compute_dt").argtypes = [C.c_void_p, C.c_void_p]
        getattr(lib, pf + "compute_dt").restype = C.c_double
        getattr(lib, pf + "worst_elem_quality").argtypes = [C.c_void_p, _i]
        getattr(lib, pf + "worst_elem_quality").restype = C.c_double
        getattr(lib, pf + "build_support").argtypes = [C.c_int, C.c_int, _i, _i, _i]
        getattr(lib, pf + "build_support").restype = C.c_int
        getattr(lib, pf + "mat_visc").argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        getattr(lib, pf + "mat_visc").restype = C.c_double
        getattr(lib, pf + "mat_rho").argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        getattr(lib, pf + "mat_rho").restype = C.c_double
        getattr(lib, pf + "mat_plastic_props").argtypes = [C.c_void_p, C.c_void_p, C.c_int,
                                                            C.c_double, _d]
        getattr(lib, pf + "bad_mesh_quality").argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, _i]
        getattr(lib, pf + "bad_mesh_quality").restype = C.c_int
        getattr(lib, pf + "sizeof_state").restype = C.c_int
        assert getattr(lib, pf + "sizeof_state")() == C.sizeof(OState)
    lib.o3_dsyevh3.argtypes = [_d, _d, _d]
    lib.o3_dsyevh3.restype = C.c_int
    lib.o3_principal_stresses3.argtypes = [_d, _d, _d]
    return lib


def ref_lib(ndims, omp=False, strict=False):
    """The reference harness.  strict: 3x3-C built without -ffast-math."""
    suffix = "_strict" if strict else ("_omp" if omp else "")
    lib = _lib(os.path.join("_ref", f"libref{ndims}d{suffix}.so"))
    pf = f"ref{ndims}_"
    getattr(lib, pf + "run").argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    getattr(lib, pf + "run").restype = C.c_double
    getattr(lib, pf + "call").argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    getattr(lib, pf + "call").restype = C.c_int
    getattr(lib, pf + "worst_quality").argtypes = [C.c_void_p, C.c_void_p, _i]
    getattr(lib, pf + "worst_quality").restype = C.c_double
    getattr(lib, pf + "nthreads").restype = C.c_int
    return lib


def ref_io_available(ndims=3):
    return os.path.exists(os.path.join(HERE, "_ref", f"libref{ndims}d_io.so"))


def ref_available(ndims=3, omp=False, strict=False):
    suffix = "_strict" if strict else ("_omp" if omp else "")
    return os.path.exists(os.path.join(HERE, "_ref", f"libref{ndims}d{suffix}.so"))


REF_CALL = {
    "compute_volume": 0, "compute_mass": 1, "compute_shape_fn": 2, "update_temperature": 3,
    "update_strain_rate": 4, "compute_dvoldt": 5, "compute_edvoldt": 6, "update_stress": 7,
    "nmd_stress": 8, "update_force": 9, "update_velocity": 10, "apply_vbcs": 11,
    "update_mesh": 12, "rotate_stress": 13, "compute_dt": 14,
}


def adjoint_lib():
    """The FP64 CPU reference gradient (des_adjoint.c)."""
    lib = _lib("liboracle_adj.so")
    lib.adj3_gradient.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _d, _d, _d]
    lib.adj3_gradient.restype = C.c_int
    lib.adj3_supported.argtypes = [C.c_void_p, C.c_void_p]
    lib.adj3_supported.restype = C.c_int
    lib.adj3_tet_grad.argtypes = [_d, _d]
    lib.adj3_tet_grad_adj.argtypes = [_d, _d, _d]
    lib.o3_sizeof_state.restype = C.c_int
    assert lib.o3_sizeof_state() == C.sizeof(OState)
    return lib


class AveragedFields:
    """Output::average_fields (output.cxx:211-251) over an OracleState stepped
    one step at a time, and the averaged arrays Output::write makes of them
    (output.cxx:83-86, 97-105, 113-131, 136-143).  Test infrastructure."""

    def __init__(self, interval):
        self.interval = interval

    def after_step(self, o):
        if o.S.steps % self.interval == 1:     # the interval's first step
            self.time0 = o.S.time
            self.coord0 = o.f["coord"].copy()
            self.strain0 = o.f["strain"].copy()
            self.stress_avg = o.f["stress"].copy()
            self.dpl_avg = o.f["delta_plstrain"].copy()
        else:
            self.stress_avg += o.f["stress"]
            self.dpl_avg += o.f["delta_plstrain"]

    def frame(self, o):
        """The arrays of an averaged frame, and the .info dt."""
        inv_dt = 1.0 / (o.S.time - self.time0)
        return {"velocity averaged": (o.f["coord"] - self.coord0) * inv_dt,
                "plastic strain-rate": self.dpl_avg * inv_dt,
                "strain-rate": (o.f["strain"] - self.strain0) * inv_dt,
                "stress averaged": self.stress_avg * (1.0 / (self.interval + 1)),
                "dt": (o.S.time - self.time0) / self.interval}


class OracleState:
    """The reference's Variables as numpy arrays plus the C view of them.

    Built from a dynearthsol_amd.model.Model; runs init() the way the
    reference does (dynearthsol.cxx:80-98) through the oracle functions.
    """

    def __init__(self, model, init=True):
        self.model = model
        self.params = model.params
        nd, npe, nstr = model.ndims, model.npe, model.nstr
        N, E = model.nnode, model.nelem
        self.nd, self.npe, self.nstr, self.N, self.E = nd, npe, nstr, N, E
        self.pf = f"o{nd}_"
        self.lib = oracle_lib()
        self.conn = np.ascontiguousarray(model.conn, dtype=np.int32)
        self.bcflag = model.bcflag
        self.csr_off = np.zeros(N + 1, dtype=np.int32)
        self.csr_ent = np.zeros(E * npe, dtype=np.int32)
        rc = getattr(self.lib, self.pf + "build_support")(
            N, E, abi.iptr(self.conn), abi.iptr(self.csr_off), abi.iptr(self.csr_ent))
        assert rc == 0
        self.f = {}
        for n in NODE_VEC:
            self.f[n] = np.zeros((N, nd))
        for n in NODE_SCAL:
            self.f[n] = np.zeros(N)
        for n in ELEM_SCAL:
            self.f[n] = np.zeros(E)
        for n in ELEM_TENS:
            self.f[n] = np.zeros((E, nstr))
        for n in ELEM_SHP:
            self.f[n] = np.zeros((E, npe))
        self.f["coord"][:] = model.coord
        self.f["coord0"][:] = model.coord
        self.ework = np.zeros(E * npe * nd + 2 * E)
        self.elem_visc = model.elem_visc
        self.elem_cohesion = model.elem_cohesion
        self.S = OState()
        S = self.S
        S.nnode, S.nelem, S.nmat = N, E, self.params.nmat
        S.markers_per_element = model.markers_per_element
        S.conn = abi.iptr(self.conn)
        S.bcflag = abi.uptr(self.bcflag)
        self.elemmarkers = model.elemmarkers
        S.elemmarkers = abi.iptr(self.elemmarkers)
        S.csr_off, S.csr_ent = abi.iptr(self.csr_off), abi.iptr(self.csr_ent)
        b, e, f = model.bfacets
        S.nbfacet = len(b)
        S.bf_bdry, S.bf_elem, S.bf_face = abi.iptr(b), abi.iptr(e), abi.iptr(f)
        S.bnormals = abi.dptr(model.bnormals)
        S.edge_vectors = abi.dptr(model.edge_vectors)
        S.edge_valid = abi.iptr(model.edge_valid)
        S.ework = abi.dptr(self.ework)
        S.elem_visc = abi.dptr(self.elem_visc)
        S.elem_cohesion = abi.dptr(self.elem_cohesion)
        self._bind()
        S.dt = S.time = 0.0
        S.steps = 0
        S.status = 0
        if init:
            self.init()

    def _bind(self):
        for n, a in self.f.items():
            setattr(self.S, n, abi.dptr(a))

    def _sync_swap(self):
        """update_mesh swaps the volume pointers in C; follow it here."""
        a = C.addressof(self.S.volume.contents)
        if a != self.f["volume"].ctypes.data:
            self.f["volume"], self.f["volume_old"] = self.f["volume_old"], self.f["volume"]
        assert C.addressof(self.S.volume.contents) == self.f["volume"].ctypes.data

    def call(self, name):
        fn = getattr(self.lib, self.pf + name)
        if name in ("compute_mass", "update_temperature", "update_force", "apply_vbcs",
                    "update_stress", "init_derived", "update_mesh"):
            fn(C.byref(self.S), C.byref(self.params))
        elif name == "compute_dt":
            self.S.dt = fn(C.byref(self.S), C.byref(self.params))
        else:
            fn(C.byref(self.S))
        self._sync_swap()
        return self.S.status

    def init(self):
        """init() (dynearthsol.cxx:80-98) then compute_dt (:316)."""
        m = self.model
        self.call("init_derived")          # temperature still zero: the reference's quirk
        self.call("apply_vbcs")
        self.f["temperature"][:] = m.temperature0
        self.f["stress"][:] = m.stress0
        self.f["strain"][:] = m.strain0
        self.f["stressyy"][:] = m.stressyy0
        self.f["plstrain"][:] = m.plstrain0
        self.S.compensation_pressure = m.compensation_pressure
        self.call("compute_dt")

    def step(self, n=1):
        rc = getattr(self.lib, self.pf + "run")(C.byref(self.S), C.byref(self.params), n)
        self._sync_swap()
        return rc

    def bad_mesh_quality(self, remeshing_option=0, max_boundary_distortion=0.25):
        i = C.c_int(0)
        r = getattr(self.lib, self.pf + "bad_mesh_quality")(
            C.byref(self.S), C.byref(self.params), remeshing_option, max_boundary_distortion, C.byref(i))
        return r, i.value

    def worst_quality(self):
        w = C.c_int()
        q = getattr(self.lib, self.pf + "worst_elem_quality")(C.byref(self.S), C.byref(w))
        return q, w.value

    def gradient(self, nsteps, coord_bar):
        """nsteps steps, then the gradient of sum(coord_bar * final coord) with
        respect to elem_visc and elem_cohesion (des_adjoint.c)."""
        lib = adjoint_lib()
        cb = np.ascontiguousarray(coord_bar, dtype=np.float64).reshape(-1)
        assert cb.size == self.N * self.nd
        vb = np.zeros(self.E)
        hb = np.zeros(self.E)
        rc = lib.adj3_gradient(C.byref(self.S), C.byref(self.params), nsteps, abi.dptr(cb),
                               abi.dptr(vb), abi.dptr(hb))
        self._sync_swap()
        if rc:
            raise ValueError("configuration not supported by the CPU adjoint")
        return vb, hb

    def set_elem_params(self, elem_visc=None, elem_cohesion=None):
        """Replace the per-element viscosity / cohesion (own copies)."""
        if elem_visc is not None:
            self.elem_visc = np.ascontiguousarray(elem_visc, dtype=np.float64).copy()
            self.S.elem_visc = abi.dptr(self.elem_visc)
        if elem_cohesion is not None:
            self.elem_cohesion = np.ascontiguousarray(elem_cohesion, dtype=np.float64).copy()
            self.S.elem_cohesion = abi.dptr(self.elem_cohesion)

    def copy(self):
        """Deep copy of the state (same mesh)."""
        o = OracleState.__new__(OracleState)
        o.__dict__.update(self.__dict__)
        o.f = {k: v.copy() for k, v in self.f.items()}
        o.ework = self.ework.copy()
        o.S = OState()
        C.pointer(o.S)[0] = self.S
        o.S.ework = abi.dptr(o.ework)
        o._bind()
        return o

    # ---- the real reference, through the harness --------------------------
    def ref_run(self, nsteps, omp=False, strict=False):
        lib = ref_lib(self.nd, omp, strict and self.nd == 3)
        t = getattr(lib, f"ref{self.nd}_run")(C.byref(self.params), C.byref(self.S), nsteps)
        return t

    def ref_call(self, name, omp=False, strict=False):
        lib = ref_lib(self.nd, omp, strict and self.nd == 3)
        rc = getattr(lib, f"ref{self.nd}_call")(C.byref(self.params), C.byref(self.S), REF_CALL[name])
        assert rc == 0
        return rc

    def ref_write_output(self, modelname, frame, segment, segflag):
        """The reference's Output::write and Output::write_checkpoint of this
        state (oracle/_ref/libref{2,3}d_io.so): <modelname>.save.<frame>,
        .chkpt.<frame> and a line of <modelname>.info."""
        lib = _lib(os.path.join("_ref", f"libref{self.nd}d_io.so"))
        fn = getattr(lib, f"ref{self.nd}_write_output")
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, C.c_int, C.c_int, _i, _i]
        fn.restype = C.c_int
        seg = np.ascontiguousarray(segment, dtype=np.int32)
        flg = np.ascontiguousarray(segflag, dtype=np.int32)
        rc = fn(C.byref(self.params), C.byref(self.S), str(modelname).encode(), frame, len(flg),
                abi.iptr(seg), abi.iptr(flg))
        assert rc == 0
        return rc

    def ref_worst_quality(self):
        lib = ref_lib(self.nd)
        w = C.c_int()
        q = getattr(lib, f"ref{self.nd}_worst_quality")(C.byref(self.params), C.byref(self.S), C.byref(w))
        return q, w.value
