"""ctypes mirror of include/graphyt_amd.h (the C ABI of the MPM step).

The structs here are byte-for-byte the C structs; `ParticleArrays` keeps the numpy
buffers a `gy_particles` view points into alive for as long as the view is used.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

GY_OK, GY_ERR_INVALID, GY_ERR_UNSUPPORTED, GY_ERR_HIP, GY_ERR_RANGE, GY_ERR_STATE = range(6)

# Boundary-condition codes — BoundaryConditions.hpp:6-77 (enum order is the ABI).
BC_NAMES = [
    "none", "grid_vx", "grid_vy", "grid_vz", "grid_temp", "grid_tflux_x", "grid_tflux_y",
    "grid_tflux_z", "grid_fx", "grid_fy", "grid_fz", "grid_vx_vy", "grid_vx_vz", "grid_vy_vz",
    "grid_fx_fy", "grid_fy_fz", "grid_fx_fz", "grid_fx_fy_fz", "grid_vx_vy_vz", "grid_vx_temp",
    "grid_vy_temp", "grid_vz_temp", "grid_fx_temp", "grid_fy_temp", "grid_fz_temp",
    "grid_vx_vy_temp", "grid_vx_vz_temp", "grid_vy_vz_temp", "grid_fx_fy_temp",
    "grid_fy_fz_temp", "grid_fx_fz_temp", "grid_fx_fy_fz_temp", "grid_vx_vy_vz_temp",
    "mp_vx", "mp_vy", "mp_vz", "mp_temp", "mp_fx", "mp_fy", "mp_fz", "mp_vx_vy", "mp_vx_vz",
    "mp_vy_vz", "mp_fx_fy", "mp_fy_fz", "mp_fx_fz", "mp_fx_fy_fz", "mp_vx_vy_vz",
    "mp_vx_temp", "mp_vy_temp", "mp_vz_temp", "mp_fx_temp", "mp_fy_temp", "mp_fz_temp",
    "mp_vx_vy_temp", "mp_vx_vz_temp", "mp_vy_vz_temp", "mp_fx_fy_temp", "mp_fy_fz_temp",
    "mp_fx_fz_temp", "mp_fx_fy_fz_temp", "mp_vx_vy_vz_temp", "mp_stress", "mp_stress_xx_yy",
    "mpstrain", "mp_vxsin", "mp_vysin", "mp_syysin",
]
BC = {name: i for i, name in enumerate(BC_NAMES)}


class GyGrid(C.Structure):
    _fields_ = [("L", C.c_double * 3), ("cellsize", C.c_double)]


class GyMaterial(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("is_rigid", C.c_int32), ("is_damage", C.c_int32),
        ("damage_model", C.c_int32), ("bulk_mod", C.c_double), ("shear_mod", C.c_double),
        ("density", C.c_double), ("viscosity", C.c_double), ("yield_stress", C.c_double),
        ("cohesion", C.c_double), ("critical_strain", C.c_double), ("phi_deg", C.c_double),
        ("psi_deg", C.c_double), ("crack_speed", C.c_double), ("damage_m", C.c_double),
        ("damage_k", C.c_double), ("thermal_conductivity", C.c_double),
        ("specific_heat", C.c_double), ("turb_stress", C.c_double),
        ("dp_apex_guard", C.c_int32), ("dp_return", C.c_int32),
    ]


class GyParams(C.Structure):
    _fields_ = [
        ("gravity", C.c_double * 3), ("damping", C.c_double), ("mu", C.c_double),
        ("cfl", C.c_double), ("psep", C.c_double), ("dt", C.c_double),
        ("is3D", C.c_int32), ("reserved", C.c_int32),
    ]


_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int32)
_PU8 = C.POINTER(C.c_uint8)


class GyParticles(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("reserved", C.c_int32),
        ("pos", _PD), ("pos0", _PD), ("vel", _PD), ("stress", _PD), ("strain", _PD),
        ("plastic_strain", _PD), ("acc_plastic_strain", _PD), ("density", _PD), ("mass", _PD),
        ("damage", _PD), ("smax", _PD), ("material", _PI), ("obj_id", _PI), ("hold", _PU8),
        ("bc_type", _PI), ("bc_val", _PD), ("volume", _PD), ("displacement", _PD),
        ("cell_id", _PI), ("temperature", _PD), ("temperature_grad", _PD),
    ]


class GyAudit(C.Structure):
    _fields_ = [
        ("mass", C.c_double), ("momentum", C.c_double * 3), ("kinetic_energy", C.c_double),
        ("strain_energy_proxy", C.c_double), ("nonfinite", C.c_int64),
        ("out_of_range", C.c_int64),
    ]


class GyTuning(C.Structure):
    """gy_tuning: launch shapes chosen by the autotuner."""
    _fields_ = [
        ("persist_wgs", C.c_int64), ("fused_waves", C.c_int32), ("grid_wgs_factor", C.c_int32),
        ("ms_per_step", C.c_double), ("from_cache", C.c_int32), ("reserved", C.c_int32),
    ]


# field name -> (components, numpy dtype)
PARTICLE_FIELDS = {
    "pos": (3, np.float64), "pos0": (3, np.float64), "vel": (3, np.float64),
    "stress": (6, np.float64), "strain": (6, np.float64), "plastic_strain": (6, np.float64),
    "acc_plastic_strain": (1, np.float64), "density": (1, np.float64), "mass": (1, np.float64),
    "damage": (1, np.float64), "smax": (1, np.float64), "material": (1, np.int32),
    "obj_id": (1, np.int32), "hold": (1, np.uint8), "bc_type": (1, np.int32),
    "bc_val": (4, np.float64), "volume": (1, np.float64), "displacement": (3, np.float64),
    "cell_id": (1, np.int32), "temperature": (1, np.float64), "temperature_grad": (3, np.float64),
}
_CTYPE = {np.float64: _PD, np.int32: _PI, np.uint8: _PU8}


@dataclass
class ParticleArrays:
    """numpy arrays behind a `gy_particles` view (shape (n, k) or (n,))."""

    n: int
    arrays: dict = field(default_factory=dict)

    @classmethod
    def empty(cls, n: int, names) -> "ParticleArrays":
        out = cls(n)
        for name in names:
            k, dt = PARTICLE_FIELDS[name]
            out.arrays[name] = np.zeros((n, k) if k > 1 else (n,), dtype=dt)
        return out

    def view(self) -> GyParticles:
        v = GyParticles()
        v.n = self.n
        for name, arr in self.arrays.items():
            k, dt = PARTICLE_FIELDS[name]
            if arr.dtype != dt or not arr.flags.c_contiguous:
                arr = np.ascontiguousarray(arr, dtype=dt)
                self.arrays[name] = arr
            if arr.shape[0] != self.n:
                raise ValueError(f"field {name} has {arr.shape[0]} rows, expected {self.n}")
            setattr(v, name, arr.ctypes.data_as(_CTYPE[dt]))
        return v

    def __getitem__(self, name):
        return self.arrays[name]


def material(model=1, bulk_mod=1e7, shear_mod=5000.0, density=1000.0, is_rigid=False,
             viscosity=1.0, yield_stress=1e7, cohesion=1e5, critical_strain=0.01, phi=40.0,
             psi=10.0, is_damage=False, damage_model=1, crack_speed=586.52, damage_m=9.0,
             damage_k=5.39e36, thermal_conductivity=391.0, specific_heat=385.2,
             turb_stress=1e10, dp_apex_guard=False, dp_return=0) -> GyMaterial:
    """A gy_material with the reference defaults (MaterialModels.cpp:7-25)."""
    m = GyMaterial()
    m.model, m.is_rigid, m.is_damage, m.damage_model = model, int(is_rigid), int(is_damage), damage_model
    m.bulk_mod, m.shear_mod, m.density, m.viscosity = bulk_mod, shear_mod, density, viscosity
    m.yield_stress, m.cohesion, m.critical_strain = yield_stress, cohesion, critical_strain
    m.phi_deg, m.psi_deg, m.crack_speed = phi, psi, crack_speed
    m.damage_m, m.damage_k = damage_m, damage_k
    m.thermal_conductivity, m.specific_heat, m.turb_stress = thermal_conductivity, specific_heat, turb_stress
    m.dp_apex_guard = int(dp_apex_guard)
    m.dp_return = int(dp_return)  # 0 closed form, 1 the reference's loop, 2 learned
    return m


def material_array(mats) -> C.Array:
    arr = (GyMaterial * len(mats))()
    for i, m in enumerate(mats):
        arr[i] = m
    return arr
