"""Binary checkpoints of a simulation state, taken from the device.

One format for both front ends: the graphyt host module (`Solver.checkpoint(path, t,
step)` / `Solver.restore(path)`, written in C++) and the C-ABI contexts of
`graphyt_amd.runtime` (`save(ctx, scene, path, t, step)` / `load(path)` here). A file is

    magic  b"GYCKPT\\x00\\x01"
    uint32 version (1), uint32 number of arrays
    float64 t, float64 dt, int64 step
    per array: uint32 name length, name (UTF-8), uint8 type (0 float64, 1 int32,
               2 uint8), uint64 element count, the elements (little-endian)

Particle arrays use the names of `abi.PARTICLE_FIELDS` (pos, pos0, vel, stress, strain,
plastic_strain, acc_plastic_strain, density, mass, volume, damage, smax, material,
obj_id, hold, bc_type, bc_val, cell_id, temperature, temperature_grad), node BCs are
node_bc_type / node_bc_val, and the grid is grid_L, grid_h, psep. `device_snapshot`
(uint8) holds the device state between two steps (gy_snapshot_save): restored into a
context of the same setup it gives back that state bit for bit, binning, node records,
tile partials and the previous step's time (read by the sine particle BCs) included.
Files without it restore from the particle arrays alone: re-binned on upload, which
reorders the FP32 sums (agreement to rounding).
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"GYCKPT\x00\x01"
VERSION = 1
_TYPES = {0: np.float64, 1: np.int32, 2: np.uint8}
_CODES = {np.dtype(np.float64): 0, np.dtype(np.int32): 1, np.dtype(np.uint8): 2}

PARTICLE_NAMES = ["pos", "pos0", "vel", "stress", "strain", "plastic_strain", "acc_plastic_strain",
                  "density", "mass", "volume", "damage", "smax", "material", "obj_id", "hold",
                  "bc_type", "bc_val", "cell_id", "temperature", "temperature_grad"]


def write(path: str, arrays: dict, t: float, step: int, dt: float = 0.0) -> None:
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<II", VERSION, len(arrays)))
        f.write(struct.pack("<ddq", float(t), float(dt), int(step)))
        for name, a in arrays.items():
            a = np.asarray(a)
            dt_ = np.dtype(a.dtype).newbyteorder("=")
            if dt_ not in _CODES:
                raise TypeError(f"{name}: dtype {a.dtype} not storable (float64, int32, uint8)")
            nb = name.encode()
            f.write(struct.pack("<I", len(nb)) + nb)
            f.write(struct.pack("<BQ", _CODES[dt_], a.size))
            f.write(np.ascontiguousarray(a, dtype=dt_.newbyteorder("<")).tobytes())


def read(path: str) -> tuple:
    """(arrays, t, step, dt); arrays are flat."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != MAGIC:
        raise ValueError(f"{path}: not a graphyt checkpoint")
    version, narr = struct.unpack_from("<II", data, 8)
    if version != VERSION:
        raise ValueError(f"{path}: checkpoint version {version}, expected {VERSION}")
    t, dt, step = struct.unpack_from("<ddq", data, 16)
    off, arrays = 40, {}
    for _ in range(narr):
        (ln,) = struct.unpack_from("<I", data, off)
        name = data[off + 4:off + 4 + ln].decode()
        off += 4 + ln
        code, count = struct.unpack_from("<BQ", data, off)
        off += 9
        dtype = np.dtype(_TYPES[code]).newbyteorder("<")
        nbytes = count * dtype.itemsize
        if off + nbytes > len(data):
            raise ValueError(f"{path}: truncated array {name}")
        arrays[name] = np.frombuffer(data, dtype, count, off).astype(dtype.newbyteorder("="))
        off += nbytes
    return arrays, t, step, dt


def save(ctx, scene, path: str, t: float, step: int) -> None:
    """Checkpoint a single-instance runtime.Context running `scene` (the scene supplies the
    grid, materials and node BCs; the particle arrays and the device snapshot come from the
    device)."""
    from .abi import PARTICLE_FIELDS
    out = ctx.download(PARTICLE_NAMES)
    arrays = {k: np.asarray(out[k]).reshape(-1) for k in PARTICLE_NAMES}
    if scene.node_bc_type is not None:
        arrays["node_bc_type"] = np.asarray(scene.node_bc_type, np.int32).reshape(-1)
        arrays["node_bc_val"] = np.asarray(scene.node_bc_val, np.float64).reshape(-1)
    arrays["grid_L"] = np.asarray(scene.L, np.float64)
    arrays["grid_h"] = np.asarray([scene.h], np.float64)
    arrays["psep"] = np.asarray([scene.psep], np.float64)
    assert set(PARTICLE_NAMES) <= set(PARTICLE_FIELDS)
    arrays["device_snapshot"] = ctx.snapshot()
    write(path, arrays, t, step, ctx.dt)


def load(path: str, scene):
    """Restore a checkpoint into `scene` (the same setup: grid, materials, parameters):
    its particles and node BCs are replaced by the checkpoint's, and its dt is set to the
    checkpoint's. Returns (scene, t, step) for `runtime.Context(scene)` and
    `iterate(t, step)`."""
    from .abi import PARTICLE_FIELDS, ParticleArrays
    arrays, t, step, dt = read(path)
    n = arrays["pos"].size // 3
    if not np.allclose(arrays["grid_L"], scene.L) or float(arrays["grid_h"][0]) != scene.h:
        raise ValueError("checkpoint grid differs from the scene's")
    pa = ParticleArrays(n)
    for k in PARTICLE_NAMES:
        if k not in arrays:  # e.g. temperature, absent from files written before it was stored
            continue
        comp, dtype = PARTICLE_FIELDS[k]
        a = arrays[k].astype(dtype)
        pa.arrays[k] = a.reshape(n, comp) if comp > 1 else a
    scene.particles = pa
    if "node_bc_type" in arrays:
        scene.node_bc_type = arrays["node_bc_type"]
        scene.node_bc_val = arrays["node_bc_val"].reshape(-1, 4)
    scene.dt = dt
    return scene, t, step


def restore(path: str, scene, graph: bool = True):
    """A runtime.Context continuing a checkpoint of `scene`'s setup: the checkpoint's
    particle arrays uploaded and, when the file holds one, its device snapshot loaded (the
    exact device state of the checkpointed step). Returns (context, t, step)."""
    from .runtime import Context
    scene, t, step = load(path, scene)
    arrays, _, _, _ = read(path)
    ctx = Context(scene, graph=graph)
    if "device_snapshot" in arrays:
        ctx.load_snapshot(arrays["device_snapshot"])
    return ctx, t, step
