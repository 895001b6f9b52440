"""Synthetic MPM workloads (scenes) with the sizes of the benchmark configurations.

A `Scene` is everything `gy_create` / `gy_set_node_bc` / `gy_upload` need: grid, material
table, run parameters, per-particle initial fields and per-node boundary conditions.
Random draws use numpy's default_rng with a recorded seed.

  * `elastic_cube`      — 3D periodic box with an elastic cube in shear-free compression
                          (0.5 m box, h = 0.02, 20^3 particles at psep = h/2, 500 steps).
  * `dp_column`         — 3D Drucker-Prager granular column collapse
                          (2x1x1 m, h = 0.01, 80x160x80 = 1,024,000 particles, 1000 steps),
                          and scaled-down versions of it for CPU parity runs.
  * `impact_block`      — elastic/perfect-plastic block impacting a fixed wall (ensemble member).
  * `dam_break_2d`      — 2D (per unit depth) Newtonian dam break in a walled box.
  * `brazilian_disk_2d` — 2D Drucker-Prager disk with Grady-Kipp damage between two rigid
                          platens.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .abi import BC, GyGrid, GyMaterial, GyParams, ParticleArrays, material

SEED = 20261203


@dataclass
class Scene:
    name: str
    L: tuple
    h: float
    psep: float
    mats: list                      # list[GyMaterial]
    gravity: tuple = (0.0, 0.0, 0.0)
    damping: float = 0.0
    mu: float = 0.0
    cfl: float = 0.1
    dt: float = 0.0                 # <= 0: the CFL rule of Solver::Solver
    steps: int = 100
    particles: ParticleArrays = None
    node_bc_type: np.ndarray = None  # int32[numNodes]
    node_bc_val: np.ndarray = None   # float64[numNodes, 4]
    meta: dict = field(default_factory=dict)
    is3D: bool = True                # False: the plane path, L[2] = 0, one node layer

    # -- grid geometry (Nodes.cpp:12, 71-73) --
    @property
    def dims(self):
        return tuple(int(1 + (self.L[k] / self.h)) for k in range(3))

    @property
    def num_nodes(self):
        L, h = self.L, self.h
        return int(round((1 + (L[0] / h)) * (1 + (L[1] / h)) * (1 + (L[2] / h))))

    @property
    def n(self):
        return self.particles.n

    def grid(self) -> GyGrid:
        g = GyGrid()
        for k in range(3):
            g.L[k] = self.L[k]
        g.cellsize = self.h
        return g

    def params(self) -> GyParams:
        p = GyParams()
        for k in range(3):
            p.gravity[k] = self.gravity[k]
        p.damping, p.mu, p.cfl, p.psep, p.dt = self.damping, self.mu, self.cfl, self.psep, self.dt
        p.is3D = 1 if self.is3D else 0
        return p

    def resolved_dt(self) -> float:
        """dt as Solver::Solver sets it (Solver.cpp:15-32), unless overridden by setDt."""
        if self.dt > 0:
            return self.dt
        maxK, minrho = 1.0, 1e15
        for m in self.mats:
            maxK = max(maxK, m.bulk_mod)
            minrho = min(minrho, m.density)
        return self.cfl * (self.h / math.sqrt(maxK / minrho))

    def node_positions(self) -> np.ndarray:
        Nx, Ny, Nz = self.dims
        k, j, i = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
        pos = np.stack([i.ravel() * self.h, j.ravel() * self.h, k.ravel() * self.h], axis=1)
        return pos


def lattice(lo, counts, psep, jitter=0.0, seed=SEED):
    """(i+1/2)*psep lattice points from corner `lo`, x fastest, optional U(-j, j)*psep jitter."""
    nx, ny, nz = counts
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([lo[0] + (i.ravel() + 0.5) * psep,
                    lo[1] + (j.ravel() + 0.5) * psep,
                    lo[2] + (k.ravel() + 0.5) * psep], axis=1)
    if jitter:
        rng = np.random.default_rng(seed)
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape) * psep
    return pos


def _particles(pos, vel=None, material=None, obj=None):
    n = pos.shape[0]
    pa = ParticleArrays(n)
    pa.arrays["pos"] = np.ascontiguousarray(pos, dtype=np.float64)
    pa.arrays["vel"] = np.zeros((n, 3)) if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
    pa.arrays["material"] = np.zeros(n, np.int32) if material is None else np.asarray(material, np.int32)
    pa.arrays["obj_id"] = np.zeros(n, np.int32) if obj is None else np.asarray(obj, np.int32)
    return pa


def elastic_cube(h=0.02, L=0.5, side=0.2, steps=500) -> Scene:
    """Periodic box with an elastic cube, v_z = -0.5 (z - z_c)/0.1, gravity -9.81 z.

    dt = 0.2 h / c_p with c_p = sqrt((K + 4G/3)/rho), set through setDt.
    """
    psep = h / 2
    n = int(round(side / psep))
    lo = (L - side) / 2
    pos = lattice((lo, lo, lo), (n, n, n), psep)
    zc = L / 2
    vel = np.zeros_like(pos)
    vel[:, 2] = -0.5 * (pos[:, 2] - zc) / 0.1
    K, G, rho = 1e6, 5e5, 1000.0
    m = material(model=1, bulk_mod=K, shear_mod=G, density=rho)
    s = Scene("elastic_cube", (L, L, L), h, psep, [m], gravity=(0, 0, -9.81),
              dt=0.2 * h / math.sqrt((K + 4 * G / 3) / rho), steps=steps)
    s.particles = _particles(pos, vel)
    s.node_bc_type = np.zeros(s.num_nodes, np.int32)
    s.node_bc_val = np.zeros((s.num_nodes, 4))
    return s


def dp_column(h=0.01, L=(2.0, 1.0, 1.0), lo=(0.03, 0.03, 0.30), size=(0.4, 0.8, 0.4),
              steps=1000, jitter=0.05, cohesion=0.0, phi=30.0, psi=0.0, seed=SEED,
              apex_guard=True, dp_return=0) -> Scene:
    """Drucker-Prager granular column collapse (model 4, rho 2000, K 5e7, G 2e7).

    Floor nodes (y <= 2h) get grid_vx_vy [0, 0]; wall nodes (x <= 2h or x >= L-2h) get
    grid_vx [0]; the floor wins where both apply. Gravity -9.81 along y; z is periodic
    because z-direction grid BCs are no-ops in the reference (Grid.cpp:872-874, 1032-1052).

    With c = 0 and psi = 0 the reference's apex return evaluates 0*inf (NaN) for every
    yielding particle in tension (ConstitutiveModels.cpp:364-365); `apex_guard` (default on)
    selects the guarded apex that returns sigma = 0 there, in the oracle and on the GPU alike.
    `dp_return` picks the multipliers of the return map (gy_material::dp_return).
    """
    psep = h / 2
    counts = tuple(int(round(size[k] / psep)) for k in range(3))
    pos = lattice(lo, counts, psep, jitter=jitter, seed=seed)
    m = material(model=4, bulk_mod=5e7, shear_mod=2e7, density=2000.0, cohesion=cohesion,
                 phi=phi, psi=psi, critical_strain=0.01, dp_apex_guard=apex_guard, dp_return=dp_return)
    s = Scene("dp_column", tuple(L), h, psep, [m], gravity=(0, -9.81, 0), steps=steps)
    s.particles = _particles(pos)
    npos = s.node_positions()
    wall = 2.0 * h
    t = np.zeros(s.num_nodes, np.int32)
    t[(npos[:, 0] <= wall) | (npos[:, 0] >= L[0] - wall)] = BC["grid_vx"]
    t[npos[:, 1] <= wall] = BC["grid_vx_vy"]
    s.node_bc_type = t
    s.node_bc_val = np.zeros((s.num_nodes, 4))
    s.meta.update(seed=seed, counts=counts)
    return s


def dp_column_small(h=0.02, steps=50, **kw) -> Scene:
    """A CPU-sized Drucker-Prager column (same material/BCs, 20x30x10 particles)."""
    return dp_column(h=h, L=(0.6, 0.5, 0.4), lo=(0.06, 0.06, 0.1), size=(0.2, 0.3, 0.1),
                     steps=steps, **kw)


def dp_column_mid(steps=1000, **kw) -> Scene:
    """The Drucker-Prager column at the benchmark resolution (h = 0.01, psep = 0.005, same
    material, BCs and dt) on a 0.6x0.5x0.3 m domain: 20x40x20 = 16,000 particles, small
    enough for the FP64 oracle to run 1000 steps in seconds."""
    return dp_column(h=0.01, L=(0.6, 0.5, 0.3), lo=(0.03, 0.03, 0.1), size=(0.1, 0.2, 0.1),
                     steps=steps, **kw)


def dp_column_4m(steps=1000, **kw) -> Scene:
    """The Drucker-Prager column at h = 0.00625 (psep = 0.003125): the same 0.4x0.8x0.4 m
    column and 2x1x1 m domain, 128x256x128 = 4,194,304 particles. As in dp_column at
    h = 0.01 (lo = 3h, 3h), the column starts 3h from the x and y origin, so the GIMP
    support of its lowest particles reaches the floor nodes (y <= 2h) at once; with
    lo = 0.03 it would fall 1.1 cm (about 12,000 steps) before touching them."""
    h = 0.00625
    return dp_column(h=h, lo=(3 * h, 3 * h, 0.30), steps=steps, **kw)


def exact_length(n_nodes: int, h: float) -> float:
    """A domain length L near (n_nodes - 1) h whose FP64 quotient L/h truncates to
    n_nodes - 1, so Nx = int(1 + L/h) (Nodes.cpp:71-73) gives n_nodes nodes."""
    L = (n_nodes - 1) * h
    while int(1 + L / h) < n_nodes:
        L = float(np.nextafter(L, np.inf))
    return L


def impact_block(h=0.01, n_nodes=24, count=20000, K=1e11, G=4e10, sy=5e8, v=100.0,
                 model=3, seed=SEED) -> Scene:
    """Perfect-plastic steel block moving at -v along x toward a fixed grid_vx wall."""
    L = exact_length(n_nodes, h)
    psep = h / 2
    side = max(1, int(round(count ** (1.0 / 3.0))))
    size = side * psep
    lo = (3 * h, (L - size) / 2, (L - size) / 2)
    pos = lattice(lo, (side, side, side), psep, jitter=0.0, seed=seed)
    vel = np.zeros_like(pos)
    vel[:, 0] = -v
    m = material(model=model, bulk_mod=K, shear_mod=G, density=7800.0, yield_stress=sy)
    s = Scene("impact_block", (L, L, L), h, psep, [m], steps=500)
    s.particles = _particles(pos, vel)
    npos = s.node_positions()
    t = np.zeros(s.num_nodes, np.int32)
    t[npos[:, 0] <= 2 * h] = BC["grid_vx"]
    s.node_bc_type = t
    s.node_bc_val = np.zeros((s.num_nodes, 4))
    return s


def impact_ensemble(count=256, steps=500, nmin=2000, nmax=64000, h=0.01, seed=SEED) -> list:
    """The batched ensemble: `count` independent perfect-plastic block impacts.

    Per instance (seeded): particle count log-uniform in [nmin, nmax] (a cube of
    round(N^(1/3))^3 lattice points at psep = h/2), grid of 16..48 nodes per axis sized to
    the block plus a random margin, rho 7800, K ~ U(1e10, 1.6e11), G ~ U(5e9, 8e10),
    yield stress ~ U(2e8, 1e9), impact speed ~ U(50, 300) m/s along -x toward a grid_vx
    wall (x <= 2h); each instance keeps its own CFL time step.
    """
    rng = np.random.default_rng(seed)
    scenes = []
    for i in range(count):
        n = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        side = max(2, int(round(n ** (1.0 / 3.0))))
        margin = int(rng.integers(0, 9))
        n_nodes = int(min(48, max(16, -(-side // 2) + 10 + margin)))
        K = float(rng.uniform(1e10, 1.6e11))
        G = float(rng.uniform(5e9, 8e10))
        sy = float(rng.uniform(2e8, 1e9))
        v = float(rng.uniform(50.0, 300.0))
        sc = impact_block(h=h, n_nodes=n_nodes, count=side ** 3, K=K, G=G, sy=sy, v=v, model=3, seed=seed + i)
        sc.name = f"impact_{i}"
        sc.steps = steps
        sc.meta.update(index=i, K=K, G=G, sy=sy, v=v)
        scenes.append(sc)
    return scenes


def gk_disk(h=0.001, R=0.05, thickness=0.02, v=0.1, steps=2000) -> Scene:
    """Brazilian-disk compression with Grady-Kipp damage (after Examples/brazillian3D.py).

    Domain 0.12 x 0.14 x 0.04 m; a disk of radius R and the given thickness (axis along
    z) at psep = h/2, object 1, elastic (model 1) with damage: rho 2600, K 3e10, G 2e10,
    crack speed 0.4 sqrt(G/rho), Weibull m 9, k 1e30. Two rigid platens (objects 2 and 3,
    material 1 = rigid with the reference's default density 1000) 10 mm wide and 3 mm
    thick touch the top and bottom of the disk and move at -v and +v along y. No gravity,
    no nodal BCs. dt by the CFL rule over both materials (min rho 1000, Solver.cpp:15-32).
    """
    psep = h / 2
    L = tuple(exact_length(int(round(x / h)) + 1, h) for x in (0.12, 0.14, 0.04))
    cx, cy, cz = 0.06, 0.07, 0.02
    nx, ny, nz = (int(round(2 * R / psep)), int(round(2 * R / psep)), int(round(thickness / psep)))
    box = lattice((cx - R, cy - R, cz - thickness / 2), (nx, ny, nz), psep)
    disk = box[(box[:, 0] - cx) ** 2 + (box[:, 1] - cy) ** 2 <= R * R]
    w, tp = 0.01, 0.003
    cnt = (int(round(w / psep)), int(round(tp / psep)), nz)
    top = lattice((cx - w / 2, cy + R, cz - thickness / 2), cnt, psep)
    bot = lattice((cx - w / 2, cy - R - tp, cz - thickness / 2), cnt, psep)
    pos = np.concatenate([disk, top, bot])
    n1, n2 = len(disk), len(top)
    obj = np.concatenate([np.full(n1, 1), np.full(n2, 2), np.full(len(bot), 3)]).astype(np.int32)
    mat = np.concatenate([np.zeros(n1), np.ones(2 * n2)]).astype(np.int32)
    vel = np.zeros_like(pos)
    vel[obj == 2, 1] = -v
    vel[obj == 3, 1] = v
    G, rho = 2e10, 2600.0
    rock = material(model=1, bulk_mod=3e10, shear_mod=G, density=rho, is_damage=True, damage_model=1,
                    crack_speed=0.4 * math.sqrt(G / rho), damage_m=9.0, damage_k=1e30)
    platen = material(is_rigid=True)
    s = Scene("gk_disk", L, h, psep, [rock, platen], steps=steps)
    s.particles = _particles(pos, vel, material=mat, obj=obj)
    s.node_bc_type = np.zeros(s.num_nodes, np.int32)
    s.node_bc_val = np.zeros((s.num_nodes, 4))
    s.meta.update(disk=n1, platen=n2)
    return s


def plane_lattice(lo, counts, psep):
    """(i+1/2)*psep lattice points in the z = 0 plane from corner `lo`, x fastest."""
    nx, ny = counts
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([lo[0] + (i.ravel() + 0.5) * psep, lo[1] + (j.ravel() + 0.5) * psep,
                     np.zeros(nx * ny)], axis=1)


def dam_break_2d(h=0.02, L=1.0, steps=500, dt=None) -> Scene:
    """2D dam break (after Example_Scripts/dambreak_2D.py): an L x L box (L[2] = 0), water
    (model 2, K 1e6, rho 1000, viscosity 1) filling [2.75h, L/2] x [2.75h, 0.75L - 2h] at
    psep = h/2, gravity -10 along y, damping 0.5, every node within 2h of a wall held by
    grid_vx_vy [0, 0]. dt = 0.2 h / c (c = sqrt(K/rho)) unless given; the script's CFL of
    1e-4 gives a dt 2000x smaller, which moves nothing in a parity run's few steps.
    """
    psep = h / 2
    lo = 2.75 * h
    counts = (int(round((0.5 * L - lo) / psep)), int(round((0.75 * L - 2 * h - lo) / psep)))
    pos = plane_lattice((lo, lo), counts, psep)
    K, rho = 1e6, 1000.0
    water = material(model=2, bulk_mod=K, shear_mod=0.0, density=rho)
    s = Scene("dam_break_2d", (L, L, 0.0), h, psep, [water], gravity=(0.0, -10.0, 0.0), damping=0.5,
              dt=0.2 * h / math.sqrt(K / rho) if dt is None else dt, steps=steps, is3D=False)
    s.particles = _particles(pos)
    npos = s.node_positions()
    wall = 2.0 * h
    t = np.zeros(s.num_nodes, np.int32)
    sel = (npos[:, 0] >= L - wall) | (npos[:, 0] <= wall) | (npos[:, 1] >= L - wall) | (npos[:, 1] <= wall)
    t[sel] = BC["grid_vx_vy"]
    s.node_bc_type = t
    s.node_bc_val = np.zeros((s.num_nodes, 4))
    return s


def brazilian_disk_2d(h=0.002, L=0.2, R=0.04, v=0.2, steps=300, jitter=0.0, seed=SEED) -> Scene:
    """2D Brazilian test (after Example_Scripts/brazillian2D.py): a disk of radius R
    (object 1) of rock, Drucker-Prager (model 4: K 3.31e9, G 2.48e9, rho 1540, c 5.2e6,
    phi = psi = 55.81, critical strain 0.1) with Grady-Kipp damage (crack speed 586.52,
    m 9, k 7e35), squeezed by two rigid platens (objects 2 and 3, 6h wide, 2.5 mm thick)
    moving at -v and +v along y. No gravity, friction 0, dt by the CFL rule (0.1).
    `jitter` moves every particle by U(-jitter, jitter)*psep in x and y: on the exact
    lattice some particles sit exactly on the GIMP support edge of a node, where the
    reference's contact (a body with vanishing mass still triggers the rigid override)
    switches discontinuously, and 1e-7 changes of the input change the result by 1e-1."""
    psep = h / 2
    c = 0.04 + R / 2.0
    n = int(round(2 * R / psep))
    box = plane_lattice((c - R, c - R), (n, n), psep)
    disk = box[(box[:, 0] - c) ** 2 + (box[:, 1] - c) ** 2 <= R * R]
    w, tp = 6 * h, 0.0025
    cnt = (int(round(w / psep)), int(round(tp / psep)))
    top = plane_lattice((c - w / 2, c + R), cnt, psep)
    bot = plane_lattice((c - w / 2, c - R - tp), cnt, psep)
    pos = np.concatenate([disk, top, bot])
    if jitter:
        pos[:, :2] += np.random.default_rng(seed).uniform(-jitter, jitter, size=(len(pos), 2)) * psep
    n1, n2 = len(disk), len(top)
    obj = np.concatenate([np.full(n1, 1), np.full(n2, 2), np.full(len(bot), 3)]).astype(np.int32)
    mat = np.concatenate([np.zeros(n1), np.ones(2 * n2)]).astype(np.int32)
    vel = np.zeros_like(pos)
    vel[obj == 2, 1] = -v
    vel[obj == 3, 1] = v
    rock = material(model=4, bulk_mod=3.31e9, shear_mod=2.48e9, density=1540.0, cohesion=5.2e6,
                    critical_strain=0.1, phi=55.81, psi=55.81, is_damage=True, damage_model=1,
                    crack_speed=586.52, damage_m=9.0, damage_k=7.0e35)
    platen = material(is_rigid=True)
    s = Scene("brazilian_disk_2d", (L, L, 0.0), h, psep, [rock, platen], cfl=0.1, steps=steps, is3D=False)
    s.particles = _particles(pos, vel, material=mat, obj=obj)
    s.node_bc_type = np.zeros(s.num_nodes, np.int32)
    s.node_bc_val = np.zeros((s.num_nodes, 4))
    s.meta.update(disk=n1, platen=n2)
    return s


def square_2d(h=0.02, L=0.5, side=0.2, vx=3.0, steps=60, mat=None, centered=False, squeeze=0.0) -> Scene:
    """2D periodic plane box with a square of side `side` near the +x face moving at vx
    with a shear v_y = 0.5 (x - x_c)/side, so it crosses the periodic wrap at L - 2h within
    tens of steps. Elastic (K 1e6, G 5e5, rho 1000) unless `mat` is given; gravity -9.81
    along y; dt = 0.2 h / c_p. `centered` puts the square in the middle of the box;
    `squeeze` adds a converging v_x = -2 squeeze (x - x_c)/side (squeeze m/s at the faces)."""
    psep = h / 2
    n = int(round(side / psep))
    lo = ((L - side) / 2 if centered else L - 2 * h - side, (L - side) / 2)
    pos = plane_lattice(lo, (n, n), psep)
    vel = np.zeros_like(pos)
    xc = lo[0] + side / 2
    vel[:, 0] = vx - 2.0 * squeeze * (pos[:, 0] - xc) / side
    vel[:, 1] = 0.5 * (pos[:, 0] - xc) / side
    K, G, rho = 1e6, 5e5, 1000.0
    m = material(model=1, bulk_mod=K, shear_mod=G, density=rho) if mat is None else mat
    s = Scene("square_2d", (L, L, 0.0), h, psep, [m], gravity=(0.0, -9.81, 0.0),
              dt=0.2 * h / math.sqrt((K + 4 * G / 3) / rho), steps=steps, is3D=False)
    s.particles = _particles(pos, vel)
    s.node_bc_type = np.zeros(s.num_nodes, np.int32)
    s.node_bc_val = np.zeros((s.num_nodes, 4))
    return s
