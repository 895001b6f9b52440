"""ctypes wrapper of the CPU oracles — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the graphyt_amd product never does.

  * kind="restatement": oracle/libcpu_ref.so, the FP64 restatement (cpu_ref.cpp)
  * kind="reference":   oracle/_ref/libgraphyt_ref.so, the reference's own sources compiled
                        in place by oracle/Makefile (only where the reference checkout exists)
  * kind="reference_gpu": oracle/_ref/libgraphyt_ref_gpu.so, the same reference build with
                        its Solver::iterate_GPU slot filled by the Graphyt-side binding of
                        the HIP step (graphyt_amd/integration/graphyt_iterate_gpu.cpp):
                        `iterate_gpu` / `synchronise` run the product through the
                        reference's own classes, the way a Graphyt script would
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from graphyt_amd.abi import GyGrid, GyMaterial, GyParams, GyParticles, ParticleArrays, material_array  # noqa: E402

LIBS = {
    "restatement": (os.environ.get("CREF_LIB", os.path.join(HERE, "libcpu_ref.so")), "cref_"),
    "reference": (os.path.join(HERE, "_ref", "libgraphyt_ref.so"), "rref_"),
    "reference_gpu": (os.path.join(HERE, "_ref", "libgraphyt_ref_gpu.so"), "rref_"),
}
_loaded = {}


def build(reference: bool | None = None) -> None:
    """Compile the oracles with oracle/Makefile (the reference only if its sources exist)."""
    target = ["all"] if reference is None else (["restatement", "reference"] if reference else ["restatement"])
    subprocess.run(["make", "-s", "-C", HERE] + target, check=True)


def available(kind: str) -> bool:
    return os.path.exists(LIBS[kind][0])


def _lib(kind: str):
    if kind in _loaded:
        return _loaded[kind]
    path, pre = LIBS[kind]
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (run oracle.build())")
    lib = C.CDLL(path)
    P = C.c_void_p
    fn = lambda name: getattr(lib, pre + name)  # noqa: E731
    fn("create").restype = P
    fn("create").argtypes = [C.POINTER(GyGrid), C.POINTER(GyMaterial), C.c_int32, C.POINTER(GyParams), C.POINTER(GyParticles)]
    fn("set_node_bc").restype = C.c_int32
    fn("set_node_bc").argtypes = [P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    fn("dt").restype = C.c_double
    fn("dt").argtypes = [P]
    fn("set_dt").argtypes = [P, C.c_double]
    fn("num_nodes").restype = C.c_int32
    fn("num_nodes").argtypes = [P]
    fn("iterate").argtypes = [P, C.c_double, C.c_int32]
    fn("iterate_n").argtypes = [P, C.c_double, C.c_int32, C.c_int32]
    fn("get").argtypes = [P, C.POINTER(GyParticles)]
    fn("get_nodes").argtypes = [P] + [C.POINTER(C.c_double)] * 5
    fn("eig3").argtypes = [C.POINTER(C.c_double)] * 3
    fn("destroy").argtypes = [P]
    if kind == "reference_gpu":
        for name in ("iterate_gpu", "synchronise", "iterate_gpu_n"):
            fn(name).restype = C.c_int32
        fn("iterate_gpu").argtypes = [P, C.c_double, C.c_int32]
        fn("iterate_gpu_n").argtypes = [P, C.c_double, C.c_int32, C.c_int32]
        fn("synchronise").argtypes = [P]
        fn("gpu_error").restype = C.c_char_p
    if kind == "restatement":
        lib.cref_constitutive.argtypes = [C.POINTER(GyMaterial)] + [C.POINTER(C.c_double)] * 5 + [C.c_double]
        lib.cref_gimp.restype = C.c_double
        lib.cref_gimp.argtypes = [C.c_double] * 4 + [C.c_int32]
    _loaded[kind] = (lib, pre)
    return _loaded[kind]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


OUT_FIELDS = ["pos", "vel", "stress", "strain", "plastic_strain", "acc_plastic_strain",
              "density", "mass", "damage", "smax", "volume", "displacement", "cell_id"]


class CpuSim:
    """One oracle simulation of a `graphyt_amd.workloads.Scene`."""

    def __init__(self, scene, kind: str = "restatement", sort_every: int | None = None,
                 tgrad_moves: bool = False, gpu_binning: bool = False):
        """Test knobs of the restatement (not the reference): sort_every re-bins every k
        steps instead of the reference's cadence (steps divisible by 10, and step 1), 1
        every step; tgrad_moves carries each particle's temperature gradient with it at a
        sort, where the reference scrambles it (cpu_ref.cpp sort_by_cell); gpu_binning
        re-bins every step by stencil-base cell with a cell's particles in object order,
        as the GPU path bins and claims body slots."""
        self.kind = kind
        if kind in ("reference", "reference_gpu"):
            # The reference's dynamic OpenMP chunk, P/(16*threads), is evaluated inside the
            # team (Solver.cpp:120-249): below 16 particles per thread it is 0 and libgomp
            # never advances the loop, so the step would hang.
            threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
            if scene.n < 16 * threads:
                raise ValueError(f"{scene.n} particles < 16 x {threads} OpenMP threads: the reference's "
                                 "chunk would be 0 and its step would hang (set OMP_NUM_THREADS lower)")
        self.lib, self.pre = _lib(kind)
        self.scene = scene
        self._mats = material_array(scene.mats)
        self._grid = scene.grid()
        self._params = scene.params()
        self._init = scene.particles.view()
        self.h = self._f("create")(C.byref(self._grid), self._mats, len(scene.mats),
                                   C.byref(self._params), C.byref(self._init))
        if not self.h:
            raise RuntimeError("oracle create failed")
        if scene.node_bc_type is not None:
            t = np.ascontiguousarray(scene.node_bc_type, np.int32)
            v = np.ascontiguousarray(scene.node_bc_val, np.float64)
            rc = self._f("set_node_bc")(self.h, len(t), t.ctypes.data_as(C.POINTER(C.c_int32)), _dp(v))
            if rc != 0:
                raise RuntimeError("oracle set_node_bc failed")
        if sort_every is not None or tgrad_moves or gpu_binning:
            if kind != "restatement":
                raise ValueError("sort_every, tgrad_moves and gpu_binning are knobs of the restatement only")
        if gpu_binning:
            self.lib.cref_set_gpu_binning.argtypes = [C.c_void_p, C.c_int32]
            self.lib.cref_set_gpu_binning(self.h, 1)
        if sort_every is not None:
            self.lib.cref_set_sort_every.argtypes = [C.c_void_p, C.c_int32]
            self.lib.cref_set_sort_every(self.h, int(sort_every))
        if tgrad_moves:
            self.lib.cref_set_tgrad_as_reference.argtypes = [C.c_void_p, C.c_int32]
            self.lib.cref_set_tgrad_as_reference(self.h, 0)

    def _f(self, name):
        return getattr(self.lib, self.pre + name)

    @property
    def dt(self) -> float:
        return self._f("dt")(self.h)

    def iterate(self, t: float, step: int) -> None:
        self._f("iterate")(self.h, t, step)

    def run(self, nsteps: int, step0: int = 0, t0: float | None = None) -> None:
        if t0 is None:
            t0 = step0 * self.dt
        self._f("iterate_n")(self.h, t0, step0, nsteps)

    # -- the reference's Solver through its iterate_GPU slot (kind "reference_gpu") --
    def _gpu_rc(self, rc):
        if rc != 0:
            raise RuntimeError(self._f("gpu_error")().decode())

    def iterate_gpu(self, t: float, step: int) -> None:
        """solver.iterate_GPU(t, step, parameters)"""
        self._gpu_rc(self._f("iterate_gpu")(self.h, t, step))

    def run_gpu(self, nsteps: int, step0: int = 0, t0: float | None = None) -> None:
        if t0 is None:
            t0 = step0 * self.dt
        self._gpu_rc(self._f("iterate_gpu_n")(self.h, t0, step0, nsteps))

    def synchronise(self) -> None:
        """solver.Synchronise(): device state back into MaterialPoints and Nodes"""
        self._gpu_rc(self._f("synchronise")(self.h))

    def get(self, fields=OUT_FIELDS) -> dict:
        pa = ParticleArrays.empty(self.scene.n, fields)
        self._f("get")(self.h, C.byref(pa.view()))
        return pa.arrays

    def nodes(self) -> dict:
        N = self._f("num_nodes")(self.h)
        out = {"mass": np.zeros(N), "vel": np.zeros((N, 3)), "vel_b1": np.zeros((N, 3)),
               "acc_b1": np.zeros((N, 3)), "force": np.zeros((N, 3))}
        self._f("get_nodes")(self.h, _dp(out["mass"]), _dp(out["vel"]), _dp(out["vel_b1"]),
                             _dp(out["acc_b1"]), _dp(out["force"]))
        return out

    def serialise(self, step: int, name: str) -> None:
        """MaterialPoints::Serialise + Nodes::Serialise of the reference (its own CSV writers)
        into output/ of the current directory (kinds "reference" and "reference_gpu")."""
        f = self._f("serialise")
        f.argtypes = [C.c_void_p, C.c_int32, C.c_char_p]
        f(self.h, int(step), name.encode())

    def close(self):
        if self.h:
            self._f("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def eig3(A, kind="restatement"):
    lib, pre = _lib(kind)
    A = np.ascontiguousarray(A, np.float64).reshape(9)
    V = np.zeros(9)
    d = np.zeros(3)
    getattr(lib, pre + "eig3")(_dp(A), _dp(V), _dp(d))
    return V.reshape(3, 3), d


def constitutive(mat, strain_rate, stress, dt, strain=None, plastic=None, epbar=0.0):
    """One constitutive update of the restatement on a single particle state."""
    lib, _ = _lib("restatement")
    er = np.ascontiguousarray(strain_rate, np.float64)
    st = np.array(stress, np.float64)
    eps = np.zeros(6) if strain is None else np.array(strain, np.float64)
    ep = np.zeros(6) if plastic is None else np.array(plastic, np.float64)
    eb = np.array([epbar], np.float64)
    lib.cref_constitutive(C.byref(mat), _dp(er), _dp(eps), _dp(st), _dp(ep), _dp(eb), dt)
    return st, ep, float(eb[0])


def gimp(xp, xn, h, lp, grad=False):
    lib, _ = _lib("restatement")
    return lib.cref_gimp(xp, xn, h, lp, int(grad))
