"""The HIP step inside the reference's own classes: Graphyt's Solver with its iterate_GPU
slot (Solver.hpp:21, stub Solver.cpp:293-302) filled by the Graphyt-side binding
(graphyt_amd/integration/graphyt_iterate_gpu.cpp), compiled from the reference's sources
with that binding into oracle/_ref/libgraphyt_ref_gpu.so. The scene is set up through
the reference's Nodes / MaterialPoints / Parameters / MaterialModels / Solver exactly as
its Python scripts do; the step loop calls solver.iterate_GPU and solver.Synchronise.

Skipped where that library was not built (it needs the reference checkout at build time).
"""
import numpy as np
import pytest

from graphyt_amd import runtime, workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ["pos", "vel", "stress", "density", "acc_plastic_strain", "mass"]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def slot(oracle_mod):
    if not oracle_mod.available("reference_gpu"):
        pytest.skip("oracle/_ref/libgraphyt_ref_gpu.so not built (needs the reference sources)")
    return oracle_mod


def report(name, a, b, keys):
    r = {k: rel(a[k], b[k]) for k in keys}
    print(f"{name}: " + ", ".join(f"{k} {v:.2e}" for k, v in r.items()), flush=True)
    return r


def test_reference_solver_iterate_gpu_matches_oracle(slot):
    """Elastic cube, 200 steps of solver.iterate_GPU in the reference's Solver: against the
    FP64 oracle at the 1e-4 bar, and equal to the same scene run through the C ABI
    directly (ctypes context) to FP32 rounding."""
    sc = W.elastic_cube(steps=200)
    ref = slot.CpuSim(sc, "reference_gpu")
    ref.run_gpu(200)
    ref.synchronise()
    a = ref.get(FIELDS)
    ref.close()
    o = slot.CpuSim(sc, "restatement")
    o.run(200)
    b = o.get(FIELDS)
    r = report("iterate_GPU vs oracle", a, b, ["pos", "vel", "stress", "density"])
    assert r["pos"] <= 1e-6 and r["vel"] <= 1e-4 and r["stress"] <= 1e-4 and r["density"] <= 1e-6
    g = runtime.Context(sc)
    g.run(200)
    c = g.download(FIELDS)
    g.close()
    r = report("iterate_GPU vs ctypes context", a, c, ["pos", "vel", "stress"])
    assert r["pos"] <= 1e-9 and r["vel"] <= 1e-6 and r["stress"] <= 1e-6


def test_reference_solver_cpu_and_gpu_steps_agree(slot):
    """The same Graphyt Solver set-up stepped with iterate_CPU (the reference's own FP64
    step) and with iterate_GPU: a cohesive Drucker-Prager column (c = 1, psi = phi, the
    reference's apex return), 100 steps."""
    sc = W.dp_column_small(cohesion=1.0, psi=30.0, apex_guard=False, steps=100)
    cpu = slot.CpuSim(sc, "reference_gpu")
    cpu.run(100)
    b = cpu.get(FIELDS)
    cpu.close()
    gpu = slot.CpuSim(sc, "reference_gpu")
    gpu.run_gpu(100)
    gpu.synchronise()
    a = gpu.get(FIELDS)
    gpu.close()
    r = report("iterate_GPU vs iterate_CPU", a, b, ["pos", "vel", "stress", "acc_plastic_strain"])
    assert r["pos"] <= 1e-6 and r["vel"] <= 1e-4 and r["stress"] <= 1e-4


def test_synchronise_is_lazy_and_resumable(slot):
    """Synchronise copies the device state into MaterialPoints and Nodes only when asked;
    stepping on after it re-uploads the host arrays and continues the same trajectory.
    (A cohesive column: Graphyt's MaterialModels has no apex guard, and the c = 0 column
    takes the reference's NaN apex return.)"""
    sc = W.dp_column_small(cohesion=1.0, psi=30.0, apex_guard=False, steps=100)
    one = slot.CpuSim(sc, "reference_gpu")
    one.run_gpu(100)
    one.synchronise()
    a = one.get(FIELDS)
    nodes = one.nodes()
    one.close()
    two = slot.CpuSim(sc, "reference_gpu")
    two.run_gpu(50)
    before = two.get(["pos"])["pos"]  # no Synchronise yet: the host arrays are the upload
    assert np.array_equal(before, sc.particles["pos"])
    two.synchronise()
    two.run_gpu(50, step0=50)
    two.synchronise()
    b = two.get(FIELDS)
    two.close()
    r = report("synchronised at step 50 vs straight", b, a, ["pos", "vel", "stress"])
    assert r["pos"] <= 1e-9 and r["vel"] <= 1e-6 and r["stress"] <= 1e-5
    # Nodes::mass after Synchronise: the P2G mass of a step. Nodes i and N-3+i hold the
    # same pair sum (Grid.cpp:320-411), so each periodic class is counted once.
    Nx, Ny, Nz = sc.dims
    k, j, i = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    canon = ((i < Nx - 3) & (j < Ny - 3) & (k < Nz - 3)).ravel()
    assert nodes["mass"][canon].sum() == pytest.approx(float(np.sum(a["mass"])), rel=1e-6)
    assert np.isfinite(nodes["vel"]).all()
