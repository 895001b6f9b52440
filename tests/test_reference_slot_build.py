"""The Graphyt-side binding (graphyt_amd/integration/graphyt_iterate_gpu.cpp) built into
the reference's own sources (oracle/_ref/libgraphyt_ref_gpu.so, oracle/Makefile): its
definitions take Graphyt's iterate_GPU / Synchronise / sync_GPU slot, every gy_* symbol it
calls is one the C ABI library exports, and the library loads without a GPU (no compute
call here; tests/test_gpu_reference_slot.py runs it)."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SLOT = os.path.join(ROOT, "oracle", "_ref", "libgraphyt_ref_gpu.so")
GYLIB = os.path.join(ROOT, "graphyt_amd", "lib", "libgraphyt_amd.so")
SYMS = {"_ZN6Solver11iterate_GPUEdiP10Parameters": "Solver::iterate_GPU(double, int, Parameters*)",
        "_ZN6Solver11SynchroniseEv": "Solver::Synchronise()",
        "_ZN6Solver8sync_GPUEv": "Solver::sync_GPU()"}


def _nm(path, *flags):
    out = subprocess.run(["nm", *flags, path], capture_output=True, text=True, check=True).stdout
    return [line.split() for line in out.splitlines() if line.strip()]


@pytest.fixture(scope="module")
def slot():
    if not os.path.exists(SLOT):
        pytest.skip("the reference build with the binding needs the reference sources at build time")
    return SLOT


def test_binding_fills_the_slot(slot):
    defined = {f[-1]: f[-2] for f in _nm(slot, "--defined-only") if len(f) >= 3}
    for sym, name in SYMS.items():
        assert defined.get(sym) == "T", f"{name} is not a strong definition in the slot build"
    assert defined.get("graphyt_gpu_release") == "T"


def test_binding_calls_only_the_c_abi(slot):
    undef = {f[-1] for f in _nm(slot, "-D", "--undefined-only")}
    exported = {f[-1] for f in _nm(GYLIB, "-D", "--defined-only")}
    gy = {s for s in undef if s.startswith("gy_")}
    assert gy >= {"gy_create", "gy_set_node_bc", "gy_upload", "gy_iterate", "gy_download",
                  "gy_download_nodes", "gy_set_dt", "gy_synchronize", "gy_destroy", "gy_last_error"}
    assert gy <= exported


def test_slot_build_loads_without_a_gpu(slot):
    lib = ctypes.CDLL(slot)
    assert hasattr(lib, "rref_iterate_gpu") and hasattr(lib, "rref_synchronise")
