"""ctypes binding of libgraphyt_amd.so and a thin Python context over the C ABI.

Loading fails loudly (ImportError) when the library has not been built; there is no
CPU fallback anywhere in the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import (GY_OK, GyAudit, GyTuning, GyGrid, GyMaterial, GyParams, GyParticles, ParticleArrays,
                  material_array)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libgraphyt_amd.so")

# Every entry point declared in include/graphyt_amd.h: name -> (restype, argtypes).
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
SIGNATURES = {
    "gy_create": (C.c_int, [_PP, C.POINTER(GyGrid), C.POINTER(GyMaterial), C.c_int32, C.POINTER(GyParams)]),
    "gy_set_node_bc": (C.c_int, [_P, C.c_int32, _I32P, _DP]),
    "gy_upload": (C.c_int, [_P, C.POINTER(GyParticles)]),
    "gy_set_dt": (C.c_int, [_P, C.c_double]),
    "gy_get_dt": (C.c_double, [_P]),
    "gy_iterate": (C.c_int, [_P, C.c_double, C.c_int32]),
    "gy_iterate_n": (C.c_int, [_P, C.c_double, C.c_int32, C.c_int32]),
    "gy_synchronize": (C.c_int, [_P]),
    "gy_download": (C.c_int, [_P, C.POINTER(GyParticles)]),
    "gy_download_nodes": (C.c_int, [_P, C.c_int32, _DP, _DP, _DP]),
    "gy_audit": (C.c_int, [_P, C.POINTER(GyAudit)]),
    "gy_debug_binning": (C.c_int, [_P, _I32P, _I32P, _I32P, _I32P, _I64P, _I64P]),
    "gy_counting_sort": (C.c_int, [_I32P, C.c_int64, C.c_int64, _I32P, _I32P, _I32P]),
    "gy_autotune_shapes": (C.c_int, [_P, C.c_int32, C.c_char_p, C.POINTER(GyTuning)]),
    "gy_get_tuning": (C.c_int, [_P, C.POINTER(GyTuning)]),
    "gy_set_tuning": (C.c_int, [_P, C.POINTER(GyTuning)]),
    "gy_snapshot_size": (C.c_int, [_P, _I64P]),
    "gy_snapshot_save": (C.c_int, [_P, _P, C.c_int64]),
    "gy_snapshot_load": (C.c_int, [_P, _P, C.c_int64]),
    # device pointers (keys, counts, offsets, order) and a hipStream_t
    "gy_counting_sort_noalloc": (C.c_int, [_P, C.c_int64, C.c_int64, _P, _P, _P, _P]),
    "gy_create_batch": (C.c_int, [_PP, C.c_int32, C.POINTER(GyGrid), C.POINTER(GyMaterial), _I32P,
                                  C.POINTER(GyParams)]),
    "gy_instance_count": (C.c_int, [_P]),
    "gy_set_node_bc_inst": (C.c_int, [_P, C.c_int32, C.c_int32, _I32P, _DP]),
    "gy_upload_batch": (C.c_int, [_P, C.c_int32, C.POINTER(GyParticles)]),
    "gy_download_inst": (C.c_int, [_P, C.c_int32, C.POINTER(GyParticles)]),
    "gy_get_dt_inst": (C.c_double, [_P, C.c_int32]),
    "gy_autotune": (C.c_int, [_P, C.c_int32, _I64P, _DP]),
    "gy_mlp_create": (C.c_int, [_PP, C.c_uint64]),
    "gy_mlp_param_count": (C.c_int, []),
    "gy_mlp_set_params": (C.c_int, [_P, _FP]),
    "gy_mlp_get_params": (C.c_int, [_P, _FP]),
    "gy_mlp_set_norm": (C.c_int, [_P, _FP, _FP, _FP, _FP, C.c_float, C.c_float, C.c_float, C.c_float]),
    "gy_mlp_forward": (C.c_int, [_P, _FP, C.c_int64, _FP]),
    "gy_mlp_gradient": (C.c_int, [_P, _FP, _FP, C.c_int64, _FP, _DP]),
    "gy_mlp_load_data": (C.c_int, [_P, _FP, _FP, C.c_int64]),
    "gy_mlp_train": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_double, C.c_double, _DP]),
    "gy_dp_map": (C.c_int, [C.POINTER(GyMaterial), _P, C.c_int32, _FP, _FP, _FP, C.c_int64, _I64P, _I64P, _DP]),
    "gy_set_return_network": (C.c_int, [_P, _P]),
    "gy_return_map_counts": (C.c_int, [_P, _I64P, _I64P, C.c_int32]),
    "gy_mlp_last_error": (C.c_char_p, [_P]),
    "gy_mlp_destroy": (None, [_P]),
    "gy_set_option": (C.c_int, [_P, C.c_int32, C.c_int64]),
    "gy_kernel_count": (C.c_int, []),
    "gy_kernel_name": (C.c_char_p, [C.c_int32]),
    "gy_kernel_stats": (C.c_int, [_P, C.c_int32, _I64P, _DP]),
    "gy_reset_kernel_stats": (C.c_int, [_P]),
    "gy_stream": (C.c_void_p, [_P]),
    "gy_last_error": (C.c_char_p, [_P]),
    "gy_version": (C.c_char_p, []),
    "gy_destroy": (None, [_P]),
}

OPT_USE_GRAPH, OPT_PROFILE, OPT_TILE, OPT_PERSIST_WGS = 1, 2, 3, 5

_lib = None


def load():
    """Load the HIP library (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m graphyt_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class GyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"graphyt_amd error {code}: {msg}")
        self.code = code


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def tuning_dict(t) -> dict:
    return {"persist_wgs": t.persist_wgs, "fused_waves": t.fused_waves, "grid_wgs_factor": t.grid_wgs_factor,
            "ms_per_step": t.ms_per_step, "from_cache": bool(t.from_cache)}


class Context:
    """One device context (gy_ctx) holding a Scene's grid, materials and particles."""

    def __init__(self, scene, graph: bool = True):
        self.lib = load()
        self.scene = scene
        self._mats = material_array(scene.mats)
        self._grid = scene.grid()
        self._params = scene.params()
        h = C.c_void_p()
        rc = self.lib.gy_create(C.byref(h), C.byref(self._grid), self._mats, len(scene.mats), C.byref(self._params))
        self.h = h
        self._check(rc)
        if scene.node_bc_type is not None:
            t = _i32(scene.node_bc_type)
            v = np.ascontiguousarray(scene.node_bc_val, np.float64)
            self._check(self.lib.gy_set_node_bc(self.h, len(t), t.ctypes.data_as(_I32P), v.ctypes.data_as(_DP)))
        self._check(self.lib.gy_upload(self.h, C.byref(scene.particles.view())))
        self.set_option(OPT_USE_GRAPH, int(graph))

    def _check(self, rc):
        if rc != GY_OK:
            msg = self.lib.gy_last_error(self.h).decode() if self.h else ""
            raise GyError(rc, msg)

    @property
    def dt(self) -> float:
        return self.lib.gy_get_dt(self.h)

    def set_dt(self, dt: float):
        self._check(self.lib.gy_set_dt(self.h, dt))

    def set_option(self, opt: int, value: int):
        self._check(self.lib.gy_set_option(self.h, opt, value))

    def iterate(self, t: float, step: int):
        self._check(self.lib.gy_iterate(self.h, t, step))

    def run(self, nsteps: int, step0: int = 0, t0: float | None = None):
        if t0 is None:
            t0 = step0 * self.dt
        self._check(self.lib.gy_iterate_n(self.h, t0, step0, nsteps))

    def synchronize(self):
        self._check(self.lib.gy_synchronize(self.h))

    def download(self, fields=("pos", "vel", "stress", "strain", "plastic_strain", "acc_plastic_strain",
                               "density", "mass", "damage", "smax", "volume", "displacement", "cell_id")):
        pa = ParticleArrays.empty(self.scene.n, fields)
        self._check(self.lib.gy_download(self.h, C.byref(pa.view())))
        return pa.arrays

    def download_nodes(self) -> dict:
        """Node mass, centre-of-mass velocity and force (m * a after BCs) of the last step."""
        nn = self.scene.num_nodes
        m, v, f = np.zeros(nn), np.zeros(3 * nn), np.zeros(3 * nn)
        self._check(self.lib.gy_download_nodes(self.h, nn, m.ctypes.data_as(_DP), v.ctypes.data_as(_DP),
                                               f.ctypes.data_as(_DP)))
        return {"mass": m, "vel": v.reshape(-1, 3), "force": f.reshape(-1, 3)}

    def autotune(self, trial_steps: int = 20):
        return autotune(self, trial_steps)

    def snapshot(self) -> np.ndarray:
        """The device state between two steps (gy_snapshot_save) as bytes."""
        n = C.c_int64()
        self._check(self.lib.gy_snapshot_size(self.h, C.byref(n)))
        buf = np.empty(n.value, np.uint8)
        self._check(self.lib.gy_snapshot_save(self.h, buf.ctypes.data_as(_P), n.value))
        return buf

    def load_snapshot(self, buf) -> None:
        """Put a snapshot of a context with the same setup back (gy_snapshot_load)."""
        buf = np.ascontiguousarray(buf, np.uint8)
        self._check(self.lib.gy_snapshot_load(self.h, buf.ctypes.data_as(_P), buf.size))

    def autotune_shapes(self, trial_steps: int = 10, cache_path: str | None = None) -> dict:
        """gy_autotune_shapes: tune (or take from `cache_path`) and apply the launch shapes."""
        t = GyTuning()
        self._check(self.lib.gy_autotune_shapes(self.h, trial_steps, cache_path.encode() if cache_path else None,
                                                C.byref(t)))
        return tuning_dict(t)

    def tuning(self) -> dict:
        t = GyTuning()
        self._check(self.lib.gy_get_tuning(self.h, C.byref(t)))
        return tuning_dict(t)

    def set_tuning(self, persist_wgs: int, fused_waves: int = 3, grid_wgs_factor: int = 2) -> None:
        t = GyTuning(persist_wgs=persist_wgs, fused_waves=fused_waves, grid_wgs_factor=grid_wgs_factor)
        self._check(self.lib.gy_set_tuning(self.h, C.byref(t)))

    def audit(self) -> dict:
        a = GyAudit()
        self._check(self.lib.gy_audit(self.h, C.byref(a)))
        return {"mass": a.mass, "momentum": tuple(a.momentum), "kinetic_energy": a.kinetic_energy,
                "strain_energy_proxy": a.strain_energy_proxy, "nonfinite": a.nonfinite,
                "out_of_range": a.out_of_range}

    def binning(self) -> dict:
        nk, ns = C.c_int64(), C.c_int64()
        self._check(self.lib.gy_debug_binning(self.h, None, None, None, None, C.byref(nk), C.byref(ns)))
        keys = np.zeros(ns.value, np.int32)
        counts = np.zeros(nk.value, np.int32)
        offsets = np.zeros(nk.value, np.int32)
        order = np.zeros(ns.value, np.int32)
        self._check(self.lib.gy_debug_binning(self.h, keys.ctypes.data_as(_I32P), counts.ctypes.data_as(_I32P),
                                              offsets.ctypes.data_as(_I32P), order.ctypes.data_as(_I32P),
                                              C.byref(nk), C.byref(ns)))
        return {"keys": keys, "counts": counts, "offsets": offsets, "order": order}

    def kernel_stats(self) -> dict:
        out = {}
        for i in range(self.lib.gy_kernel_count()):
            n, ms = C.c_int64(), C.c_double()
            self._check(self.lib.gy_kernel_stats(self.h, i, C.byref(n), C.byref(ms)))
            if n.value:
                out[self.lib.gy_kernel_name(i).decode()] = (n.value, ms.value)
        return out

    def reset_kernel_stats(self):
        self._check(self.lib.gy_reset_kernel_stats(self.h))

    @property
    def stream(self) -> int:
        return self.lib.gy_stream(self.h)

    def set_return_network(self, net: "Mlp") -> None:
        """gy_set_return_network: the network of materials with dp_return = learned."""
        self._check(self.lib.gy_set_return_network(self.h, net.h))

    def return_map_counts(self, reset: bool = False) -> dict:
        """Learned-material samples through the network and through the closed-form fallback."""
        a, b = C.c_int64(), C.c_int64()
        self._check(self.lib.gy_return_map_counts(self.h, C.byref(a), C.byref(b), int(reset)))
        return {"network": a.value, "fallback": b.value}

    def close(self):
        if self.h:
            self.lib.gy_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchContext:
    """Many independent scenes advanced together (gy_create_batch / gy_upload_batch):
    every step call advances all of them in the same launches."""

    def __init__(self, scenes, graph: bool = True):
        self.lib = load()
        self.scenes = list(scenes)
        k = len(self.scenes)
        grids = (GyGrid * k)(*[sc.grid() for sc in self.scenes])
        params = (GyParams * k)(*[sc.params() for sc in self.scenes])
        mats = [m for sc in self.scenes for m in sc.mats]
        self._mats = material_array(mats)
        nmat = np.array([len(sc.mats) for sc in self.scenes], np.int32)
        h = C.c_void_p()
        rc = self.lib.gy_create_batch(C.byref(h), k, grids, self._mats, nmat.ctypes.data_as(_I32P), params)
        self.h = h
        self._check(rc)
        for i, sc in enumerate(self.scenes):
            if sc.node_bc_type is not None:
                t = _i32(sc.node_bc_type)
                v = np.ascontiguousarray(sc.node_bc_val, np.float64)
                self._check(self.lib.gy_set_node_bc_inst(self.h, i, len(t), t.ctypes.data_as(_I32P),
                                                         v.ctypes.data_as(_DP)))
        self._views = (GyParticles * k)(*[sc.particles.view() for sc in self.scenes])
        self._check(self.lib.gy_upload_batch(self.h, k, self._views))
        self.set_option(OPT_USE_GRAPH, int(graph))

    _check = Context._check
    set_option = Context.set_option
    run = Context.run
    iterate = Context.iterate
    synchronize = Context.synchronize
    audit = Context.audit
    kernel_stats = Context.kernel_stats
    reset_kernel_stats = Context.reset_kernel_stats
    close = Context.close
    __del__ = Context.__del__

    @property
    def n(self) -> int:
        return sum(sc.n for sc in self.scenes)

    @property
    def dt(self) -> float:
        return self.lib.gy_get_dt(self.h)

    def dt_of(self, i: int) -> float:
        return self.lib.gy_get_dt_inst(self.h, i)

    def download(self, i: int, fields=("pos", "vel", "stress", "plastic_strain", "acc_plastic_strain", "density")):
        pa = ParticleArrays.empty(self.scenes[i].n, fields)
        self._check(self.lib.gy_download_inst(self.h, i, C.byref(pa.view())))
        return pa.arrays

    def autotune_shapes(self, trial_steps: int = 10, cache_path: str | None = None) -> dict:
        return autotune_shapes(self, trial_steps, cache_path)

    def autotune(self, trial_steps: int = 20):
        return autotune(self, trial_steps)


def autotune_shapes(ctx, trial_steps: int = 10, cache_path: str | None = None) -> dict:
    """gy_autotune_shapes on any context (single or batched)."""
    t = GyTuning()
    ctx._check(ctx.lib.gy_autotune_shapes(ctx.h, trial_steps, cache_path.encode() if cache_path else None,
                                          C.byref(t)))
    return tuning_dict(t)


def autotune(ctx, trial_steps: int = 20) -> dict:
    """gy_autotune: pick the persistent-kernel workgroup count by timing trial steps; the
    particle state is restored afterwards."""
    wgs, ms = C.c_int64(), C.c_double()
    ctx._check(ctx.lib.gy_autotune(ctx.h, trial_steps, C.byref(wgs), C.byref(ms)))
    return {"persist_wgs": wgs.value, "ms_per_step": ms.value}


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class Mlp:
    """The return-map network on the device (gy_mlp_*): 14 -> 64 -> 64 -> 64 -> 1, SiLU,
    BF16 matrix-core products with FP32 accumulation."""
    N_IN, N_OUT = 14, 1

    def __init__(self, seed: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.gy_mlp_create(C.byref(h), seed)
        self.h = h
        self._check(rc)
        self.nparam = self.lib.gy_mlp_param_count()

    def _check(self, rc):
        if rc != GY_OK:
            msg = self.lib.gy_mlp_last_error(self.h).decode() if self.h else ""
            raise GyError(rc, msg)

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(self.nparam, np.float32)
        self._check(self.lib.gy_mlp_get_params(self.h, p.ctypes.data_as(_FP)))
        return p

    @params.setter
    def params(self, p):
        p = _f32(p)
        assert p.size == self.nparam
        self._check(self.lib.gy_mlp_set_params(self.h, p.ctypes.data_as(_FP)))

    def set_norm(self, in_mean, in_std, in_lo, in_hi, out_scale=1.0, out_offset=0.0, out_lo=-1e30, out_hi=1e30):
        a = [_f32(x) for x in (in_mean, in_std, in_lo, in_hi)]
        self._norm = a + [float(out_scale), float(out_offset), float(out_lo), float(out_hi)]
        self._check(self.lib.gy_mlp_set_norm(self.h, *[x.ctypes.data_as(_FP) for x in a], float(out_scale),
                                             float(out_offset), float(out_lo), float(out_hi)))

    def forward(self, x) -> np.ndarray:
        x = _f32(x)
        y = np.zeros((x.shape[0], self.N_OUT), np.float32)
        self._check(self.lib.gy_mlp_forward(self.h, x.ctypes.data_as(_FP), x.shape[0], y.ctypes.data_as(_FP)))
        return y

    def gradient(self, x, t):
        x, t = _f32(x), _f32(t)
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_double()
        self._check(self.lib.gy_mlp_gradient(self.h, x.ctypes.data_as(_FP), t.ctypes.data_as(_FP), x.shape[0],
                                             g.ctypes.data_as(_FP), C.byref(loss)))
        return g, loss.value

    def train(self, x, t, epochs=10, batch=4096, lr=1e-3, lr_decay=0.8) -> float:
        x, t = _f32(x), _f32(t)
        self._check(self.lib.gy_mlp_load_data(self.h, x.ctypes.data_as(_FP), t.ctypes.data_as(_FP), x.shape[0]))
        loss = C.c_double()
        self._check(self.lib.gy_mlp_train(self.h, epochs, batch, lr, lr_decay, C.byref(loss)))
        return loss.value

    def close(self):
        if self.h:
            self.lib.gy_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


DP_CLOSED, DP_REFERENCE, DP_LEARNED = 0, 1, 2   # GY_DP_RETURN_*


def dp_map(mat, samples, net: "Mlp | None" = None, mode: "int | None" = None) -> dict:
    """The Drucker-Prager return map over raw samples [n, 13] on the GPU (gy_dp_map): the
    closed form (default), the reference's loop (mode DP_REFERENCE) or, given `net`, the
    gated network. Returns out [n, 7], the cone multipliers, counts and the kernel time."""
    lib = load()
    if mode is None:
        mode = DP_LEARNED if net is not None else DP_CLOSED
    x = _f32(samples)
    y = np.zeros((x.shape[0], 7), np.float32)
    mult = np.zeros(x.shape[0], np.float32)
    nn_, nf, ms = C.c_int64(), C.c_int64(), C.c_double()
    rc = lib.gy_dp_map(C.byref(mat), net.h if net else None, mode, x.ctypes.data_as(_FP), y.ctypes.data_as(_FP),
                       mult.ctypes.data_as(_FP), x.shape[0], C.byref(nn_), C.byref(nf), C.byref(ms))
    if rc != GY_OK:
        raise GyError(rc, lib.gy_mlp_last_error(net.h).decode() if net else "gy_dp_map failed")
    return {"out": y, "multiplier": mult, "network": nn_.value, "fallback": nf.value, "ms": ms.value}


def counting_sort(keys, nkeys: int):
    """Stable counting sort of int32 keys in [0, nkeys) on the GPU (gy_counting_sort)."""
    lib = load()
    keys = _i32(keys)
    counts = np.zeros(nkeys, np.int32)
    offsets = np.zeros(nkeys, np.int32)
    order = np.zeros(len(keys), np.int32)
    rc = lib.gy_counting_sort(keys.ctypes.data_as(_I32P), len(keys), nkeys, counts.ctypes.data_as(_I32P),
                              offsets.ctypes.data_as(_I32P), order.ctypes.data_as(_I32P))
    if rc != GY_OK:
        raise GyError(rc, "gy_counting_sort failed")
    return counts, offsets, order
