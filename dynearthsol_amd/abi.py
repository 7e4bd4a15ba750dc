"""ctypes mirror of the C-ABI structs declared in include/des_amd.h.

Pure data definitions: importing this module loads no library.
"""
import ctypes as C

DES_MAX_MAT = 8
DES_NBDRY = 10

DES_OK = 0
DES_ERR_INPUT = 1
DES_ERR_IO = 2
DES_ERR_MESH = 10
DES_ERR_RUNTIME = 11
DES_ERR_ASSERT = 12
DES_ERR_HIP = 100
DES_ERR_STATE = 101

# rheology bit flags (matprops.hpp:47-53)
RH_ELASTIC = 1
RH_VISCOUS = 2
RH_PLASTIC = 4
RH_MAXWELL = 3
RH_EP = 5
RH_EVP = 7

DES_FP64 = 0
DES_MIXED = 1

_M = C.c_double * DES_MAX_MAT


class DesParams(C.Structure):
    _fields_ = [
        ("ndims", C.c_int),
        ("rheol_type", C.c_int),
        ("nmat", C.c_int),
        ("is_plane_strain", C.c_int),
        ("visc_min", C.c_double), ("visc_max", C.c_double),
        ("tension_max", C.c_double), ("therm_diff_max", C.c_double),
        ("rho0", _M), ("alpha", _M),
        ("bulk_modulus", _M), ("shear_modulus", _M),
        ("visc_exponent", _M), ("visc_coefficient", _M),
        ("visc_activation_energy", _M),
        ("heat_capacity", _M), ("therm_cond", _M),
        ("radiogenic_heating", _M),
        ("pls0", _M), ("pls1", _M),
        ("cohesion0", _M), ("cohesion1", _M),
        ("friction_angle0", _M), ("friction_angle1", _M),
        ("dilation_angle0", _M), ("dilation_angle1", _M),
        ("gravity", C.c_double), ("characteristic_speed", C.c_double),
        ("inertial_scaling", C.c_double), ("dt_fraction", C.c_double),
        ("fixed_dt", C.c_double),
        ("damping_factor", C.c_double),
        ("damping_option", C.c_int), ("ref_pressure_option", C.c_int),
        ("is_quasi_static", C.c_int), ("has_thermal_diffusion", C.c_int),
        ("surface_process_option", C.c_int),
        ("surface_temperature", C.c_double), ("mantle_temperature", C.c_double),
        ("has_fixed_bottom_temperature", C.c_int),
        ("has_winkler_foundation", C.c_int),
        ("winkler_delta_rho", C.c_double),
        ("has_elastic_foundation", C.c_int),
        ("elastic_foundation_constant", C.c_double),
        ("has_water_loading", C.c_int),
        ("vbc_types", C.c_int * DES_NBDRY),
        ("vbc_values", C.c_double * DES_NBDRY),
        ("zlength", C.c_double),
        ("resolution", C.c_double), ("smallest_size", C.c_double),
        ("min_quality", C.c_double),
        ("max_vbc_val", C.c_double),
    ]


_P = C.POINTER
_d = _P(C.c_double)
_i = _P(C.c_int)


class DesMesh(C.Structure):
    _fields_ = [
        ("nnode", C.c_int), ("nelem", C.c_int),
        ("connectivity", _i),
        ("coord", _d),
        ("bcflag", _P(C.c_uint)),
        ("nbfacet", C.c_int),
        ("bfacet_bdry", _i), ("bfacet_elem", _i), ("bfacet_face", _i),
        ("bnormals", _d),
        ("edge_vectors", _d),
        ("edge_valid", _i),
        ("elemmarkers", _i),
        ("markers_per_element", C.c_int),
    ]


FIELD_NAMES = (
    "coord", "coord0", "vel", "force",
    "temperature", "mass", "tmass", "volume_n", "ntmp",
    "volume", "volume_old", "edvoldt", "plstrain", "delta_plstrain",
    "dpressure", "stressyy",
    "strain_rate", "strain", "stress",
    "shpdx", "shpdy", "shpdz",
    "elem_visc", "elem_cohesion",
)


class DesFields(C.Structure):
    _fields_ = [(n, _d) for n in FIELD_NAMES]


def dptr(a):
    """double* into a C-contiguous float64 numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype.name == "float64" and a.flags.c_contiguous
    return a.ctypes.data_as(_d)


def iptr(a):
    if a is None:
        return None
    assert a.dtype.name == "int32" and a.flags.c_contiguous
    return a.ctypes.data_as(_i)


def uptr(a):
    if a is None:
        return None
    assert a.dtype.name == "uint32" and a.flags.c_contiguous
    return a.ctypes.data_as(_P(C.c_uint))


class DesNamedArray(C.Structure):
    """des_named_array: a named block of bytes for des_write_checkpoint."""
    _fields_ = [("name", C.c_char_p), ("data", C.c_void_p), ("bytes", C.c_size_t)]
