"""Checkpoint and restart on the GPU, and VTK output from a running solver.

A checkpoint holds the particle arrays and the device snapshot (gy_snapshot_save): a
fresh context of the same setup restored from it holds the checkpointed device state bit
for bit (downloads and a second snapshot equal the first byte for byte). Its continuation
follows the uninterrupted run as closely as two uninterrupted runs follow each other: the
step is reproducible to rounding (the order of its atomic adds varies), stated per test.
The restart in the middle of the particle-BC scene checks the previous step's time, which
the sine BCs read (TimeIntegration.cpp:240-265), survives the restart."""
import os
import sys

import numpy as np
import pytest

from graphyt_amd import checkpoint, runtime, vtk, workloads as W

from test_gpu_graphyt_api import build_cube, rel

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import scenes  # noqa: E402

pytestmark = pytest.mark.gpu

FIELDS = ["pos", "vel", "stress", "strain", "plastic_strain", "acc_plastic_strain", "density", "cell_id"]


def test_snapshot_restores_the_device_state_bit_for_bit(tmp_path):
    sc = W.dp_column_small(steps=60)
    g = runtime.Context(sc)
    g.run(30)
    f = str(tmp_path / "col.ckpt")
    checkpoint.save(g, sc, f, 30 * g.dt, 30)
    snap = g.snapshot()
    a = g.download(FIELDS)
    g.run(30, step0=30)
    a2 = g.download(FIELDS)
    g.close()
    g2, t, s0 = checkpoint.restore(f, W.dp_column_small(steps=60))
    assert s0 == 30 and t == pytest.approx(30 * g2.dt, rel=1e-15)
    b = g2.download(FIELDS)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(g2.snapshot(), snap)
    g2.run(30, step0=s0, t0=t)
    b2 = g2.download(FIELDS)
    g2.close()
    # two uninterrupted runs of the same 60 steps, for the reproducibility of the step
    g3 = runtime.Context(sc)
    g3.run(60)
    c2 = g3.download(FIELDS)
    g3.close()
    for k in ("pos", "vel", "stress"):
        print(f"{k}: restored vs original {rel(b2[k], a2[k]):.2e}, two runs {rel(c2[k], a2[k]):.2e}", flush=True)
    assert rel(b2["pos"], a2["pos"]) <= 1e-9
    assert rel(b2["vel"], a2["vel"]) <= 1e-6
    assert rel(b2["stress"], a2["stress"]) <= 1e-5


def test_restart_keeps_the_time_of_the_sine_bcs(tmp_path):
    """The particle-BC scene (sine velocity and stress BCs read the previous step's time)
    restarted at step 20 against the uninterrupted 40 steps; the same restart from the
    particle arrays alone (no snapshot: a fresh context knows no previous time) misses
    by a whole step of prescribed velocity."""
    sc, steps, t_of = scenes.build("bcs")
    g = runtime.Context(sc)
    for s in range(20):
        g.iterate(t_of(s), s)
    f = str(tmp_path / "bcs.ckpt")
    checkpoint.save(g, sc, f, t_of(20), 20)
    for s in range(20, steps):
        g.iterate(t_of(s), s)
    a = g.download(["pos", "vel", "stress"])
    g.close()
    sc2, _, _ = scenes.build("bcs")
    g2, t, s0 = checkpoint.restore(f, sc2)
    for s in range(s0, steps):
        g2.iterate(t_of(s), s)
    b = g2.download(["pos", "vel", "stress"])
    g2.close()
    sine = (sc.particles["bc_type"] >= 65)
    assert sine.sum() == 3
    print("restart with snapshot: pos %.2e vel %.2e" % (rel(b["pos"], a["pos"]), rel(b["vel"], a["vel"])), flush=True)
    assert rel(b["pos"], a["pos"]) <= 1e-9 and rel(b["vel"], a["vel"]) <= 1e-6
    assert np.allclose(b["pos"][sine], a["pos"][sine], rtol=0, atol=1e-10)
    # without the snapshot
    sc3, _, _ = scenes.build("bcs")
    sc3, t, s0 = checkpoint.load(f, sc3)
    g3 = runtime.Context(sc3)
    for s in range(s0, steps):
        g3.iterate(t_of(s), s)
    c = g3.download(["pos", "vel", "stress"])
    g3.close()
    miss = np.abs(c["pos"][sine] - a["pos"][sine]).max()
    print(f"restart from the arrays alone: sine-BC particles off by {miss:.2e} m", flush=True)
    assert miss > 1e-7


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
    r = {k: rel(v, a[k]) for k, v in (("pos", mp2.getPosArr()), ("vel", mp2.getVelArr()),
                                       ("stress", mp2.getStressArr()), ("disp", mp2.getDispArr()))}
    print("host module restart:", ", ".join(f"{k} {v:.2e}" for k, v in r.items()), flush=True)
    assert r["pos"] <= 1e-9 and r["vel"] <= 1e-6 and r["stress"] <= 1e-5 and r["disp"] <= 1e-6


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
