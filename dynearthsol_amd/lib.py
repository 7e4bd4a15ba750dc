"""ctypes binding of libdes_amd.so (the C-ABI in include/des_amd.h).

Loading fails loudly when the library is not built or no HIP device is
usable: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdes_amd.so")

EXPORTS = (
    "des_create", "des_destroy", "des_last_error", "des_upload_fields",
    "des_download_fields", "des_set_scalars", "des_get_scalars",
    "des_compute_volume", "des_compute_mass", "des_compute_shape_fn",
    "des_init_derived", "des_update_temperature", "des_update_strain_rate",
    "des_compute_dvoldt", "des_compute_edvoldt", "des_update_stress",
    "des_nmd_stress", "des_update_force", "des_update_velocity",
    "des_apply_vbcs", "des_update_mesh", "des_rotate_stress",
    "des_compute_dt", "des_worst_quality", "des_step", "des_run", "des_run_timed",
    "des_sync", "des_bad_mesh_quality", "des_upload_elemmarkers", "des_write_frame",
    "des_get_tuning", "des_batch_create", "des_batch_destroy", "des_batch_last_error",
    "des_batch_upload_fields", "des_batch_download_fields", "des_batch_set_scalars",
    "des_batch_get_scalars", "des_batch_init_derived", "des_batch_apply_vbcs",
    "des_batch_compute_dt", "des_batch_run", "des_batch_sync", "des_batch_worst_quality",
    "des_batch_tuning", "des_adj_forward", "des_adj_backward", "des_adj_release", "des_set_tuning",
    "des_set_mesh_order", "des_set_averaging", "des_average_fields", "des_write_frame_averaged",
    "des_worst_quality_noalloc", "des_get_tuning_all", "des_set_tuning_all", "des_batch_tuning_all",
    "des_write_checkpoint", "des_remesh_finish", "des_upload_fields_device", "des_download_fields_device",
    "des_adj_backward_device", "des_set_sweep_order", "des_set_launch_mode", "des_batch_set_launch_mode",
    "des_set_force_layout",
)

NTUNE = 6
TUNE_MAXV = 6
TUNE_PASSES = ("pass_a", "pass_c", "pass_d", "pass_e", "pass_f", "pass_b")
_NODE_VARIANTS = ("block256-depth4", "block128-depth4", "block64-depth4", "block256-depth8", "block256-depth16")
TUNE_MEANING = {
    "pass_a": ("lds-staged", "direct", "direct waves>=6"), "pass_e": ("lds-staged", "direct", "lds-staged waves>=8"),
    "pass_c": ("waves>=1", "waves>=2", "waves>=3", "waves>=4", "lds-staged waves>=1", "lds-staged waves>=2"),
    "pass_d": _NODE_VARIANTS, "pass_f": _NODE_VARIANTS, "pass_b": ("depth4", "depth8", "depth16"),
}

PER_FUNCTION = (
    "compute_volume", "compute_mass", "compute_shape_fn", "init_derived",
    "update_temperature", "update_strain_rate", "compute_dvoldt",
    "compute_edvoldt", "update_stress", "nmd_stress", "update_force",
    "update_velocity", "apply_vbcs", "update_mesh", "rotate_stress",
)

_lib = None


class DesError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"des status {code}: {msg}")
        self.code = code


def load():
    """The library handle with argument types set."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run python -m dynearthsol_amd.build")
    # One HIP runtime per process.  torch's wheel carries its own
    # libamdhip64 / libhsa-runtime64; a second runtime in the same process
    # cannot open the GPU once the first has (the GPU's address space is
    # held by one runtime's device file), so whichever initialises second
    # sees no device.  torch's copies answer to the sonames this library
    # needs (libamdhip64.so.7, libhsa-runtime64.so.1): with torch loaded
    # first the dynamic linker binds this library to torch's runtime, and
    # device pointers pass between the two.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.des_create.argtypes = [C.POINTER(abi.DesParams), C.POINTER(abi.DesMesh), C.POINTER(vp)]
    lib.des_create.restype = C.c_int
    lib.des_destroy.argtypes = [vp]
    lib.des_destroy.restype = None
    lib.des_last_error.argtypes = [vp]
    lib.des_last_error.restype = C.c_char_p
    lib.des_upload_fields.argtypes = [vp, C.POINTER(abi.DesFields)]
    lib.des_download_fields.argtypes = [vp, C.POINTER(abi.DesFields)]
    lib.des_set_scalars.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_double]
    lib.des_get_scalars.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int)]
    for f in PER_FUNCTION + ("step", "sync"):
        getattr(lib, "des_" + f).argtypes = [vp]
        getattr(lib, "des_" + f).restype = C.c_int
    lib.des_compute_dt.argtypes = [vp, C.POINTER(C.c_double)]
    lib.des_worst_quality.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.des_worst_quality_noalloc.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.des_run.argtypes = [vp, C.c_int, C.c_int]
    lib.des_run_timed.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.des_bad_mesh_quality.argtypes = [vp, C.c_int, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.des_upload_elemmarkers.argtypes = [vp, C.POINTER(C.c_int)]
    lib.des_write_frame.argtypes = [vp, C.c_char_p, C.c_int, C.c_double]
    lib.des_write_frame_averaged.argtypes = [vp, C.c_char_p, C.c_int, C.c_double]
    lib.des_remesh_finish.argtypes = [vp, C.c_int]
    lib.des_upload_fields_device.argtypes = [vp, C.POINTER(abi.DesFields)]
    lib.des_download_fields_device.argtypes = [vp, C.POINTER(abi.DesFields)]
    lib.des_adj_backward_device.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.des_write_checkpoint.argtypes = [vp, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.c_int, C.POINTER(abi.DesNamedArray)]
    lib.des_set_averaging.argtypes = [vp, C.c_int]
    lib.des_average_fields.argtypes = [vp]
    lib.des_get_tuning.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.des_set_tuning.argtypes = [vp, C.c_int]
    lib.des_get_tuning_all.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.des_set_tuning_all.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    lib.des_batch_tuning_all.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.des_set_mesh_order.argtypes = [C.c_int]
    lib.des_set_mesh_order.restype = None
    lib.des_set_sweep_order.argtypes = [C.c_int]
    lib.des_set_launch_mode.argtypes = [vp, C.c_int]
    lib.des_set_force_layout.argtypes = [C.c_int]
    lib.des_set_force_layout.restype = None
    lib.des_batch_set_launch_mode.argtypes = [vp, C.c_int]
    lib.des_set_sweep_order.restype = None
    lib.des_batch_create.argtypes = [C.c_int, C.POINTER(abi.DesParams), C.POINTER(abi.DesMesh), C.POINTER(vp)]
    lib.des_batch_destroy.argtypes = [vp]
    lib.des_batch_destroy.restype = None
    lib.des_batch_last_error.argtypes = [vp]
    lib.des_batch_last_error.restype = C.c_char_p
    lib.des_batch_upload_fields.argtypes = [vp, C.c_int, C.POINTER(abi.DesFields)]
    lib.des_batch_download_fields.argtypes = [vp, C.c_int, C.POINTER(abi.DesFields)]
    lib.des_batch_set_scalars.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double]
    lib.des_batch_get_scalars.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int)]
    for f in ("init_derived", "apply_vbcs", "sync"):
        getattr(lib, "des_batch_" + f).argtypes = [vp]
    lib.des_batch_compute_dt.argtypes = [vp, C.POINTER(C.c_double)]
    lib.des_batch_run.argtypes = [vp, C.c_int, C.c_int]
    lib.des_batch_worst_quality.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.des_batch_tuning.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.des_adj_forward.argtypes = [vp, C.c_int, C.c_int]
    lib.des_adj_backward.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.des_adj_release.argtypes = [vp]
    lib.des_adj_release.restype = None
    _lib = lib
    return lib


def set_mesh_order(mode):
    """Device numbering of meshes created from now on in this process:
    1 Hilbert-curve order (default), 0 the host's order.  Results do not depend on it."""
    load().des_set_mesh_order(int(mode))


def set_sweep_order(mode):
    """Block order of the fused node passes, process-wide: 1 (default) each
    XCD walks passes B, D and F from the top of its range down, 0 upwards.
    Results do not depend on it."""
    load().des_set_sweep_order(int(mode))


def set_force_layout(mode):
    """Force contributions of the fused FP64 schedule, process-wide: 0
    node-major CSR slots addressed by list start + one-byte position
    (default), 1 element-major records, 2 node-major slots addressed by their
    full positions (0 and 2 also for the mixed schedule's FP32 slots).
    Results do not depend on it."""
    load().des_set_force_layout(int(mode))


class DesModel:
    """One model on the current HIP device, set up like the reference's init()
    (dynearthsol.cxx:80-98, 316) from a dynearthsol_amd.model.Model."""

    def __init__(self, model, init=True):
        self.lib = load()
        self.model = model
        self.nd, self.npe, self.nstr = model.ndims, model.npe, model.nstr
        self.N, self.E = model.nnode, model.nelem
        self.params = model.params
        self._mesh = model.mesh_struct()
        ctx = C.c_void_p()
        rc = self.lib.des_create(C.byref(self.params), C.byref(self._mesh), C.byref(ctx))
        self.ctx = ctx
        if rc:
            msg = self.lib.des_last_error(ctx).decode() if ctx.value else "des_create failed"
            if ctx.value:
                self.lib.des_destroy(ctx)
                self.ctx = C.c_void_p()
            raise DesError(rc, msg)
        if init:
            self.init()

    def close(self):
        if self.ctx and self.ctx.value:
            self.lib.des_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise DesError(rc, self.lib.des_last_error(self.ctx).decode())

    def call(self, name):
        self._check(getattr(self.lib, "des_" + name)(self.ctx))

    def upload(self, **arrays):
        f, keep = _fields_struct(arrays)
        self._check(self.lib.des_upload_fields(self.ctx, C.byref(f)))

    def init(self):
        m = self.model
        self.call("init_derived")          # temperature is still zero: the reference's quirk
        self.call("apply_vbcs")
        self.upload(temperature=m.temperature0, stress=m.stress0, strain=m.strain0,
                    stressyy=m.stressyy0, plstrain=m.plstrain0, elem_visc=m.elem_visc,
                    elem_cohesion=m.elem_cohesion)
        self._check(self.lib.des_set_scalars(self.ctx, 0.0, 0.0, 0, m.compensation_pressure))
        self.compute_dt()

    def compute_dt(self):
        dt = C.c_double()
        self._check(self.lib.des_compute_dt(self.ctx, C.byref(dt)))
        return dt.value

    def step(self, n=1):
        for _ in range(n):
            self._check(self.lib.des_step(self.ctx))

    def run(self, n, precision=abi.DES_FP64):
        self._check(self.lib.des_run(self.ctx, n, precision))

    def set_launch_mode(self, mode):
        """0: one launch per fused pass (default); 1: one persistent launch per run."""
        self._check(self.lib.des_set_launch_mode(self.ctx, int(mode)))

    PASS_NAMES = ("pass_a", "pass_b", "pass_c", "pass_d", "pass_e", "facet", "pass_f", "dt", "pass_c_ret")

    def run_timed(self, n, precision=abi.DES_FP64):
        """des_run with per-pass HIP-event timing: {pass: (total_ms, launches)}."""
        ms = (C.c_double * 10)()
        cnt = (C.c_int * 10)()
        self._check(self.lib.des_run_timed(self.ctx, n, precision, ms, cnt))
        return {self.PASS_NAMES[k]: (ms[k], cnt[k]) for k in range(len(self.PASS_NAMES)) if cnt[k]}

    def set_tuning(self, variant):
        self._check(self.lib.des_set_tuning(self.ctx, variant))

    def sync(self):
        self._check(self.lib.des_sync(self.ctx))

    def scalars(self):
        dt, t, st = C.c_double(), C.c_double(), C.c_int()
        self._check(self.lib.des_get_scalars(self.ctx, C.byref(dt), C.byref(t), C.byref(st)))
        return dt.value, t.value, st.value

    def bad_mesh_quality(self, remeshing_option=0, max_boundary_distortion=0.25):
        """(reason, index) as bad_mesh_quality (remeshing.cxx:1882-1934)."""
        r, i = C.c_int(), C.c_int()
        self._check(self.lib.des_bad_mesh_quality(self.ctx, remeshing_option, max_boundary_distortion,
                                                  C.byref(r), C.byref(i)))
        return r.value, i.value

    def write_frame(self, modelname, frame, run_time=0.0):
        """<modelname>.save.<frame> and a line of <modelname>.info (output.cxx:79-208)."""
        self._check(self.lib.des_write_frame(self.ctx, str(modelname).encode(), frame, run_time))

    def remesh_finish(self, remeshing_option=0):
        """remesh()'s tail on the device (des_remesh_finish)."""
        self._check(self.lib.des_remesh_finish(self.ctx, remeshing_option))

    def write_checkpoint(self, modelname, frame, segment, segflag, extra=()):
        """<modelname>.chkpt.<frame> (output.cxx:254-276): the host's boundary
        segments (nseg x ndims, int) and flags (nseg), then the named byte
        blocks of extra ((name, bytes) pairs: the host's marker sets)."""
        seg = np.ascontiguousarray(segment, dtype=np.int32).reshape(-1, self.nd)
        flg = np.ascontiguousarray(segflag, dtype=np.int32).reshape(-1)
        assert len(seg) == len(flg)
        blocks = [(str(n).encode(), bytes(b)) for n, b in extra]
        arr = (abi.DesNamedArray * max(len(blocks), 1))()
        keep = []
        for k, (n, b) in enumerate(blocks):
            buf = C.create_string_buffer(b, len(b))
            keep.append(buf)
            arr[k].name, arr[k].data, arr[k].bytes = n, C.cast(buf, C.c_void_p), len(b)
        self._check(self.lib.des_write_checkpoint(self.ctx, str(modelname).encode(), frame, len(flg),
                                                  abi.iptr(seg), abi.iptr(flg), len(blocks), arr))

    def write_frame_averaged(self, modelname, frame, run_time=0.0):
        """A regular frame with the averaged fields (output.cxx:79-208)."""
        self._check(self.lib.des_write_frame_averaged(self.ctx, str(modelname).encode(), frame, run_time))

    def set_averaging(self, average_interval):
        """Output::average_fields after every step of des_run (0: off)."""
        self._check(self.lib.des_set_averaging(self.ctx, average_interval))

    def average_fields(self):
        """Output::average_fields for the step just completed (per-function drivers)."""
        self._check(self.lib.des_average_fields(self.ctx))

    def upload_elemmarkers(self, counts):
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        assert counts.shape == (self.E, self.params.nmat)
        self._check(self.lib.des_upload_elemmarkers(self.ctx, abi.iptr(counts)))

    def worst_quality(self, noalloc=False):
        """worst_elem_quality (geometry.cxx:499-512): (quality, host element);
        noalloc: through des_worst_quality_noalloc (no device allocation)."""
        q, w = C.c_double(), C.c_int()
        fn = self.lib.des_worst_quality_noalloc if noalloc else self.lib.des_worst_quality
        self._check(fn(self.ctx, C.byref(q), C.byref(w)))
        return q.value, w.value

    def reset(self):
        """Back to the model's initial state (coordinates, fields, scalars),
        keeping the context and its per-element parameters."""
        m = self.model
        z = np.zeros((self.N, self.nd))
        self.upload(coord=m.coord, coord0=m.coord, vel=z, force=z,
                    delta_plstrain=np.zeros(self.E), dpressure=np.zeros(self.E))
        self.init()

    def set_elem_params(self, elem_visc=None, elem_cohesion=None):
        """Upload per-element viscosity and / or cohesion."""
        self.upload(elem_visc=elem_visc, elem_cohesion=elem_cohesion)

    def adj_forward(self, nsteps, window=100):
        """nsteps steps of the differentiated run, checkpointing every window
        steps, or keeping every step's records when window is 0
        (des_adj_forward)."""
        self._check(self.lib.des_adj_forward(self.ctx, nsteps, window))

    def adj_backward(self, coord_bar):
        """Gradients of the loss whose gradient with respect to the final
        coordinates is coord_bar (nnode x ndims): (d/d elem_visc, d/d elem_cohesion)."""
        cb = np.ascontiguousarray(coord_bar, dtype=np.float64)
        assert cb.shape == (self.N, self.nd)
        vb = np.zeros(self.E)
        hb = np.zeros(self.E)
        self._check(self.lib.des_adj_backward(self.ctx, abi.dptr(cb), abi.dptr(vb), abi.dptr(hb)))
        return vb, hb

    def upload_device(self, **ptrs):
        """des_upload_fields_device: fields given as device addresses (int)
        of arrays in the host's row order and AoS layout."""
        f = abi.DesFields()
        for k, v in ptrs.items():
            if v is not None:
                setattr(f, k, C.cast(C.c_void_p(int(v)), C.POINTER(C.c_double)))
        self._check(self.lib.des_upload_fields_device(self.ctx, C.byref(f)))

    def download_device(self, **ptrs):
        """des_download_fields_device into device addresses (int)."""
        f = abi.DesFields()
        for k, v in ptrs.items():
            if v is not None:
                setattr(f, k, C.cast(C.c_void_p(int(v)), C.POINTER(C.c_double)))
        self._check(self.lib.des_download_fields_device(self.ctx, C.byref(f)))

    def adj_backward_device(self, coord_bar, visc_bar, cohesion_bar):
        """des_adj_backward_device on device addresses (int): coord_bar
        (nnode x ndims) in, visc_bar and cohesion_bar (nelem) out."""
        self._check(self.lib.des_adj_backward_device(self.ctx, C.c_void_p(int(coord_bar)),
                                                     C.c_void_p(int(visc_bar)), C.c_void_p(int(cohesion_bar))))

    def adj_release(self):
        self.lib.des_adj_release(self.ctx)

    def tuning_all(self, precision=abi.DES_FP64):
        """{pass: (variant or -1, [mean ms per variant])} of the autotuner."""
        v = (C.c_int * NTUNE)()
        ms = (C.c_double * (TUNE_MAXV * NTUNE))()
        self._check(self.lib.des_get_tuning_all(self.ctx, precision, v, ms))
        return {p: (v[k], list(ms[TUNE_MAXV * k:TUNE_MAXV * (k + 1)])) for k, p in enumerate(TUNE_PASSES)}

    def set_tuning_all(self, precision, choices):
        """Fix the autotuner's choices ({pass: variant} or a sequence in
        TUNE_PASSES order; -1 lets the next long run tune that pass)."""
        if isinstance(choices, dict):
            choices = [choices[p][0] if isinstance(choices[p], tuple) else choices[p] for p in TUNE_PASSES]
        self._check(self.lib.des_set_tuning_all(self.ctx, precision, (C.c_int * NTUNE)(*choices)))

    def tuning(self):
        """(pass-C variant chosen by the autotuner or -1, [ms per variant])."""
        v = C.c_int()
        ms = (C.c_double * TUNE_MAXV)()
        self._check(self.lib.des_get_tuning(self.ctx, C.byref(v), ms))
        return v.value, list(ms)

    def download(self, names=None):
        """Fields as numpy arrays in the reference's AoS layout."""
        out, f = _field_buffers(self.model, names)
        self._check(self.lib.des_download_fields(self.ctx, C.byref(f)))
        self.sync()
        return out


def _field_shapes(model):
    nd, npe, nstr, N, E = model.ndims, model.npe, model.nstr, model.nnode, model.nelem
    shapes = {
        "coord": (N, nd), "coord0": (N, nd), "vel": (N, nd), "force": (N, nd),
        "temperature": (N,), "mass": (N,), "tmass": (N,), "volume_n": (N,), "ntmp": (N,),
        "volume": (E,), "volume_old": (E,), "edvoldt": (E,), "plstrain": (E,),
        "delta_plstrain": (E,), "dpressure": (E,), "stressyy": (E,),
        "strain_rate": (E, nstr), "strain": (E, nstr), "stress": (E, nstr),
        "shpdx": (E, npe), "shpdz": (E, npe),
    }
    if nd == 3:
        shapes["shpdy"] = (E, npe)
    return shapes


def _field_buffers(model, names):
    shapes = _field_shapes(model)
    if names is None:
        names = list(shapes)
    out = {n: np.zeros(shapes[n]) for n in names}
    f = abi.DesFields()
    for n, a in out.items():
        setattr(f, n, abi.dptr(a))
    return out, f


def _fields_struct(arrays):
    f = abi.DesFields()
    keep = []
    for k, a in arrays.items():
        if a is None:
            continue
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        setattr(f, k, abi.dptr(a))
    return f, keep


class DesBatch:
    """Many dynearthsol_amd.model.Model instances stepping together in one
    batched context per dimension (des_batch_* in include/des_amd.h); each
    model is set up like DesModel.init."""

    def __init__(self, models, init=True):
        self.lib = load()
        self.models = list(models)
        n = len(self.models)
        self._params = (abi.DesParams * n)(*[m.params for m in self.models])
        self._mstructs = [m.mesh_struct() for m in self.models]
        self._meshes = (abi.DesMesh * n)(*self._mstructs)
        b = C.c_void_p()
        rc = self.lib.des_batch_create(n, self._params, self._meshes, C.byref(b))
        self.b = b
        if rc:
            msg = self.lib.des_batch_last_error(b).decode() if b.value else "des_batch_create failed"
            self.close()
            raise DesError(rc, msg)
        if init:
            self.init()

    def close(self):
        if self.b and self.b.value:
            self.lib.des_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.models)

    def _check(self, rc):
        if rc:
            raise DesError(rc, self.lib.des_batch_last_error(self.b).decode())

    def upload(self, k, **arrays):
        f, keep = _fields_struct(arrays)
        self._check(self.lib.des_batch_upload_fields(self.b, k, C.byref(f)))

    def init(self):
        self._check(self.lib.des_batch_init_derived(self.b))
        self._check(self.lib.des_batch_apply_vbcs(self.b))
        for k, m in enumerate(self.models):
            self.upload(k, temperature=m.temperature0, stress=m.stress0, strain=m.strain0,
                        stressyy=m.stressyy0, plstrain=m.plstrain0, elem_visc=m.elem_visc,
                        elem_cohesion=m.elem_cohesion)
            self._check(self.lib.des_batch_set_scalars(self.b, k, 0.0, 0.0, 0, m.compensation_pressure))
        return self.compute_dt()

    def compute_dt(self):
        dt = np.zeros(len(self.models))
        self._check(self.lib.des_batch_compute_dt(self.b, abi.dptr(dt)))
        return dt

    def run(self, n, precision=abi.DES_FP64):
        self._check(self.lib.des_batch_run(self.b, n, precision))

    def set_launch_mode(self, mode):
        """0: one launch per fused pass per group (default); 1: every model in one persistent launch."""
        self._check(self.lib.des_batch_set_launch_mode(self.b, int(mode)))

    def sync(self):
        self._check(self.lib.des_batch_sync(self.b))

    def scalars(self, k):
        dt, t, st = C.c_double(), C.c_double(), C.c_int()
        self._check(self.lib.des_batch_get_scalars(self.b, k, C.byref(dt), C.byref(t), C.byref(st)))
        return dt.value, t.value, st.value

    def worst_quality(self):
        q = np.zeros(len(self.models))
        w = np.zeros(len(self.models), dtype=np.int32)
        self._check(self.lib.des_batch_worst_quality(self.b, abi.dptr(q), abi.iptr(w)))
        return q, w

    def tuning(self, ndims):
        v = C.c_int()
        ms = (C.c_double * TUNE_MAXV)()
        self._check(self.lib.des_batch_tuning(self.b, ndims, C.byref(v), ms))
        return v.value, list(ms)

    def tuning_all(self, ndims, precision=abi.DES_FP64):
        """{pass: (variant or -1, [mean ms per variant])} of the group of that dimension."""
        v = (C.c_int * NTUNE)()
        ms = (C.c_double * (TUNE_MAXV * NTUNE))()
        self._check(self.lib.des_batch_tuning_all(self.b, ndims, precision, v, ms))
        return {p: (v[k], list(ms[TUNE_MAXV * k:TUNE_MAXV * (k + 1)])) for k, p in enumerate(TUNE_PASSES)}

    def download(self, k, names=None):
        out, f = _field_buffers(self.models[k], names)
        self._check(self.lib.des_batch_download_fields(self.b, k, C.byref(f)))
        self.sync()
        return out
