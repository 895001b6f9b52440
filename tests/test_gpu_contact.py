"""Two-body contact on the GPU (several objects or rigid materials): body-slot claims,
the rigid override, the contact correction with and without friction, against the FP64
oracle on the same seeded inputs (relative L2 over all particles)."""
import numpy as np
import pytest

from graphyt_amd import runtime, workloads as W
from graphyt_amd.abi import material

pytestmark = pytest.mark.gpu

F = ["pos", "vel", "stress"]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_both(oracle_mod, sc, steps, graph=False):
    g = runtime.Context(sc, graph=graph)
    g.run(steps)
    a = g.download(F)
    audit = g.audit()
    g.close()
    o = oracle_mod.CpuSim(sc, "restatement")
    o.run(steps)
    return a, o.get(F), audit


def platen_scene(steps):
    """A deformable block pressed by a rigid platen (objects 1 and 2; Solver.cpp:42-56
    gives rigid particles unit mass)."""
    sc = W.elastic_cube(steps=steps)
    sc.gravity = (0.0, 0.0, 0.0)
    pos = sc.particles["pos"]
    top = pos[:, 2] > np.quantile(pos[:, 2], 0.85)
    vel = np.zeros_like(pos)
    vel[top, 2] = -0.3
    sc.mats = [material(model=1, bulk_mod=1e6, shear_mod=5e5, density=1000.0), material(is_rigid=True)]
    sc.particles.arrays.update(material=np.where(top, 1, 0).astype(np.int32),
                               obj_id=np.where(top, 2, 1).astype(np.int32), vel=vel)
    return sc


def two_blocks(steps, mu=0.0, shear=0.0):
    """Two halves of the cube as separate deformable objects moving into each other (and
    sliding along y with `shear`)."""
    sc = W.elastic_cube(steps=steps)
    sc.gravity = (0.0, 0.0, 0.0)
    sc.mu = mu
    pos = sc.particles["pos"]
    left = pos[:, 0] < 0.25
    vel = np.zeros_like(pos)
    vel[left, 0], vel[~left, 0] = 0.5, -0.5
    vel[left, 1], vel[~left, 1] = shear, -shear
    sc.particles.arrays.update(obj_id=np.where(left, 1, 2).astype(np.int32), vel=vel)
    return sc


def test_rigid_platen_200_steps(oracle_mod):
    a, b, audit = run_both(oracle_mod, platen_scene(200), 200)
    for k, tol in (("pos", 1e-6), ("vel", 1e-4), ("stress", 1e-4)):
        print(f"{k}: rel L2 {rel(a[k], b[k]):.2e} (tol {tol:.0e})", flush=True)
        assert rel(a[k], b[k]) <= tol, k
    assert audit["nonfinite"] == 0


def test_two_deformable_bodies_in_contact(oracle_mod):
    a, b, _ = run_both(oracle_mod, two_blocks(100), 100)
    for k, tol in (("pos", 1e-6), ("vel", 1e-4), ("stress", 1e-4)):
        print(f"{k}: rel L2 {rel(a[k], b[k]):.2e} (tol {tol:.0e})", flush=True)
        assert rel(a[k], b[k]) <= tol, k


def test_contact_with_friction(oracle_mod):
    a, b, _ = run_both(oracle_mod, two_blocks(80, mu=0.3, shear=0.2), 80)
    for k, tol in (("pos", 1e-6), ("vel", 1e-4), ("stress", 1e-4)):
        print(f"{k}: rel L2 {rel(a[k], b[k]):.2e} (tol {tol:.0e})", flush=True)
        assert rel(a[k], b[k]) <= tol, k


def test_contact_graph_replay_matches_eager():
    outs = []
    for graph in (False, True):
        g = runtime.Context(platen_scene(60), graph=graph)
        g.run(60)
        outs.append(g.download(F))
        g.close()
    for k, tol in (("pos", 1e-7), ("vel", 1e-5), ("stress", 1e-5)):
        assert rel(outs[1][k], outs[0][k]) <= tol, k


def test_brazilian_disk_damage_growth(oracle_mod):
    """The same disk with the platens at 20 m/s for 600 steps: the contact load takes the
    strain past the Grady-Kipp activation strain and damage grows in ~2400 of the 19,960
    particles (FP64 eig3 of every stressed particle, DamageModels.cpp). Reached on MI355X:
    the same 2376 particles damaged as in the oracle, damage 1.2e-6, stress 4.3e-7, s_max
    9.1e-7, velocity 6.3e-7 relative L2."""
    sc = W.gk_disk(h=0.004, v=20.0, steps=600)
    g = runtime.Context(sc)
    g.run(600)
    a = g.download(["pos", "vel", "stress", "damage", "smax"])
    g.close()
    o = oracle_mod.CpuSim(sc, "restatement")
    o.run(600)
    b = o.get(["pos", "vel", "stress", "damage", "smax"])
    grown = int((b["damage"] > 0).sum())
    print(f"damaged particles: oracle {grown}, GPU {int((a['damage'] > 0).sum())}; "
          f"max damage {b['damage'].max():.3f}", flush=True)
    assert grown > 1000
    out = {k: rel(a[k], b[k]) for k in ("pos", "vel", "stress", "damage", "smax")}
    print(", ".join(f"{k} {v:.2e}" for k, v in out.items()), flush=True)
    for k, tol in (("pos", 1e-6), ("vel", 1e-4), ("stress", 1e-4), ("damage", 1e-4), ("smax", 1e-4)):
        assert out[k] <= tol, k


def test_brazilian_disk_full_size_first_steps(oracle_mod):
    """The contact disk at its full size (1,266,720 particles, h = 0.001), 30 steps, against
    the oracle binning and claiming as the GPU does. The exact lattice puts particles on
    GIMP support edges, where the contact's rigid override switches on vanishing body mass:
    perturbing the initial positions by 1e-9 (relative) moves the FP64 oracle's own disk
    velocity by 25 % and stress by 35 % after these 30 steps. The GPU stays within 1e-2
    (reached on MI355X: velocity 1.6e-3, stress 2.4e-3 over the disk particles)."""
    sc = W.gk_disk()
    g = runtime.Context(sc)
    g.run(30)
    a = g.download(["pos", "vel", "stress"])
    au = g.audit()
    g.close()
    o = oracle_mod.CpuSim(sc, "restatement", gpu_binning=True)
    o.run(30)
    b = o.get(["pos", "vel", "stress"])
    o.close()
    disk = sc.particles.arrays["obj_id"] == 1
    out = {k: rel(a[k][disk], b[k][disk]) for k in ("pos", "vel", "stress")}
    print(", ".join(f"{k} {v:.2e}" for k, v in out.items()), flush=True)
    assert au["nonfinite"] == 0
    assert out["pos"] <= 1e-6 and out["vel"] <= 1e-2 and out["stress"] <= 1e-2


def test_brazilian_disk_with_rigid_platens(oracle_mod):
    """The Grady-Kipp disk between two rigid platens (three objects), scaled to h = 0.004
    (19,960 particles), 300 steps: contact, rigid override and the FP64 eig3 path.
    Velocity, stress and s_max follow the oracle to 1e-4 (reached on MI355X: 2.0e-5,
    1.7e-5, 5.1e-5)."""
    sc = W.gk_disk(h=0.004, steps=300)
    g = runtime.Context(sc)
    g.run(300)
    a = g.download(["pos", "vel", "stress", "smax"])
    g.close()
    o = oracle_mod.CpuSim(sc, "restatement")
    o.run(300)
    b = o.get(["pos", "vel", "stress", "smax"])
    assert (b["smax"] > 0).any()
    for k, tol in (("pos", 1e-6), ("vel", 1e-4), ("stress", 1e-4), ("smax", 1e-4)):
        print(f"{k}: rel L2 {rel(a[k], b[k]):.2e} (tol {tol:.0e})", flush=True)
        assert rel(a[k], b[k]) <= tol, k
