"""The C-ABI library loads and exports every entry point include/des_amd.h
declares (no device call is made)."""
import ctypes as C
import os
import re

import pytest

from dynearthsol_amd import abi, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "des_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(des_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_reference_functions():
    names = _declared()
    for f in ("des_update_temperature", "des_update_strain_rate", "des_compute_dvoldt",
              "des_compute_edvoldt", "des_update_stress", "des_nmd_stress", "des_update_force",
              "des_update_velocity", "des_apply_vbcs", "des_update_mesh", "des_rotate_stress",
              "des_compute_dt", "des_worst_quality", "des_run", "des_step"):
        assert f in names


@pytest.mark.skipif(not os.path.exists(lib.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    h = lib.load()
    missing = [n for n in _declared() if not hasattr(h, n)]
    assert not missing, missing
    assert set(lib.EXPORTS) <= set(_declared())


def test_struct_layouts_match_header():
    """Sizes of the ctypes mirrors against the header's field list."""
    text = open(os.path.join(ROOT, "include", "des_amd.h")).read()
    body = text[text.index("typedef struct des_fields"):text.index("} des_fields;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    n_ptr = sum(len(re.findall(r"\*\s*\w+", line)) for line in body.splitlines()
                if line.strip().startswith("double"))
    assert n_ptr == len(abi.FIELD_NAMES) == len(abi.DesFields._fields_)
    assert C.sizeof(abi.DesFields) == 8 * n_ptr


def test_library_and_torch_share_one_hip_runtime():
    """Loaded through the package, the library binds to the HIP runtime
    torch carries: one libamdhip64 and one libhsa-runtime64 mapped in the
    process (a second runtime could not open the GPU the first holds)."""
    pytest.importorskip("torch")
    lib.load()
    import torch  # noqa: F401
    maps = open("/proc/self/maps").read().split("\n")
    for stem in ("libamdhip64", "libhsa-runtime64"):
        paths = {ln.split()[-1] for ln in maps if stem in ln}
        assert len(paths) == 1, paths
