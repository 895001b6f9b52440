"""Checkpoint and restart on the GPU, and VTK output from a running solver: a run that
checkpoints at step 40 and goes on to step 80 against a fresh setup restored from that
checkpoint and advanced over the same 40 steps. The restored run is re-binned on upload,
which reorders FP32 sums only, so the two agree to rounding (relative L2 stated per
field)."""
import numpy as np
import pytest

from graphyt_amd import checkpoint, runtime, vtk, workloads as W

from test_gpu_graphyt_api import build_cube, rel

pytestmark = pytest.mark.gpu


def test_host_module_checkpoint_restart(tmp_path):
    sc = W.elastic_cube(steps=80)
    nodes, mp, prm, solver = build_cube(sc)
    for step in range(40):
        solver.iterate_GPU(step * solver.dt, step, prm)
    f = str(tmp_path / "cube.ckpt")
    solver.checkpoint(f, 40 * solver.dt, 40)
    for step in range(40, 80):
        solver.iterate_GPU(step * solver.dt, step, prm)
    a = {k: getattr(mp, g)().copy() for k, g in (("pos", "getPosArr"), ("vel", "getVelArr"),
                                                  ("stress", "getStressArr"), ("disp", "getDispArr"))}
    nodes2, mp2, prm2, solver2 = build_cube(sc)
    t, s0 = solver2.restore(f)
    assert s0 == 40 and t == pytest.approx(40 * solver.dt, rel=1e-15)
    for step in range(s0, 80):
        solver2.iterate_GPU(step * solver2.dt, step, prm2)
    assert rel(mp2.getPosArr(), a["pos"]) <= 1e-7
    assert rel(mp2.getVelArr(), a["vel"]) <= 1e-5
    assert rel(mp2.getStressArr(), a["stress"]) <= 1e-5
    assert rel(mp2.getDispArr(), a["disp"]) <= 1e-5


def test_context_checkpoint_restart(tmp_path):
    sc = W.dp_column_small(steps=60)
    g = runtime.Context(sc)
    g.run(30)
    f = str(tmp_path / "col.ckpt")
    checkpoint.save(g, sc, f, 30 * g.dt, 30)
    g.run(30, step0=30)
    a = g.download(["pos", "vel", "stress", "acc_plastic_strain"])
    g.close()
    sc2, t, s0 = checkpoint.load(f, W.dp_column_small(steps=60))
    g2 = runtime.Context(sc2)
    g2.run(30, step0=s0, t0=t)
    b = g2.download(["pos", "vel", "stress", "acc_plastic_strain"])
    g2.close()
    assert rel(b["pos"], a["pos"]) <= 1e-7
    assert rel(b["vel"], a["vel"]) <= 1e-4
    assert rel(b["stress"], a["stress"]) <= 1e-3


def test_vtp_from_a_running_solver_writes_the_last_step(tmp_path):
    """The writer holds the getter arrays; after more steps Write fetches the device state
    itself (no Synchronise call in the loop, as in the reference's scripts)."""
    sc = W.elastic_cube(steps=20)
    nodes, mp, prm, solver = build_cube(sc)
    w = vtk.VTPWriter(mp.numParticles)
    w.AddPositions("Position", mp.getPosArr(), 3, vtk.Order.ijk)
    w.AddArray("Velocity", mp.getVelArr(), 3, vtk.Order.ijk)
    vti = vtk.VTIWriter([nodes.Nx, nodes.Ny, nodes.Nz], [0.0, 0.0, 0.0], nodes.cellsize)
    vti.AddArray("Mass", nodes.getMassArr(), 1, vtk.Order.ijk)
    for step in range(20):
        solver.iterate_GPU(step * solver.dt, step, prm)
    w.Write(str(tmp_path / "p.vtp"))
    vti.Write(str(tmp_path / "g.vti"))
    r, rg = vtk.read(str(tmp_path / "p.vtp")), vtk.read(str(tmp_path / "g.vti"))
    solver.Synchronise()
    assert np.array_equal(r["Position"], mp.getPosArr()) and np.array_equal(r["Velocity"], mp.getVelArr())
    assert not np.array_equal(r["Position"], sc.particles["pos"].reshape(-1))
    assert rg["Mass"].sum() == pytest.approx(mp.getMassArr().sum(), rel=1e-5)
