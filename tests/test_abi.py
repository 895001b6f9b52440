"""The C-ABI library builds, loads, and exports every entry point include/graphyt_amd.h
declares; the ctypes mirror matches the header's struct layouts. No compute calls."""
import ctypes as C
import os
import re

import pytest

from graphyt_amd import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "graphyt_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gy_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_step_entry_points():
    names = declared_functions()
    for required in ("gy_create", "gy_upload", "gy_iterate", "gy_iterate_n", "gy_download",
                     "gy_audit", "gy_counting_sort", "gy_destroy"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = runtime.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers exactly the declared set
    assert sorted(runtime.SIGNATURES) == declared_functions()


def test_version_string():
    assert b"gfx950" in runtime.load().gy_version()


def test_struct_sizes_match_header_layout():
    # sizes implied by the C declarations (x86-64 SysV alignment)
    assert C.sizeof(abi.GyGrid) == 32
    assert C.sizeof(abi.GyMaterial) == 16 + 15 * 8 + 8
    assert C.sizeof(abi.GyParams) == 8 * 8 + 8
    assert C.sizeof(abi.GyParticles) == 8 + 21 * 8
    assert C.sizeof(abi.GyAudit) == 6 * 8 + 2 * 8
    assert C.sizeof(abi.GyTuning) == 8 + 4 + 4 + 8 + 4 + 4


def test_bc_codes_match_reference_enum():
    # BoundaryConditions.hpp:6-77: 68 values, none = 0 ... mp_syysin = 67
    assert len(abi.BC_NAMES) == 68
    assert abi.BC["grid_vx_vy"] == 11 and abi.BC["grid_vx_vy_vz"] == 18
    assert abi.BC["mp_vx"] == 33 and abi.BC["mp_stress_xx_yy"] == 63 and abi.BC["mp_syysin"] == 67
    text = open(HEADER).read()
    for name, val in (("GY_BC_GRID_VX_VY", 11), ("GY_BC_MP_FX_FY", 43), ("GY_BC_MP_SYYSIN", 67)):
        assert re.search(rf"{name}\s*=\s*{val}\b", text)


def test_create_without_gpu_fails_cleanly():
    """On a machine with no GPU the context cannot be created; the error is reported, no
    CPU path is taken."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from graphyt_amd import runtime, workloads\n"
            "try:\n    runtime.Context(workloads.elastic_cube())\n    print('GPU')\n"
            "except runtime.GyError as e:\n    print('ERR', e.code)\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120).stdout
    assert out.strip() in ("GPU", "ERR 3")
