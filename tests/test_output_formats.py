"""Output formats: the VTK XML writers (graphyt_amd.vtk, the calls of the reference's
scripts, template_script.py:134-181) and the binary checkpoint shared by the graphyt
host module and graphyt_amd.checkpoint. No GPU calls."""
import numpy as np
import pytest

from graphyt_amd import checkpoint, graphyt, vtk


def small_setup(n_side=4):
    nodes = graphyt.Nodes([0.2, 0.2, 0.2], 0.02)
    mp = graphyt.MaterialPoints(nodes)
    mp.psep = 0.01
    lo = 0.05
    i = 0
    for a in range(n_side):
        for b in range(n_side):
            for c in range(n_side):
                mp.add(i, lo + 0.01 * a, lo + 0.01 * b, lo + 0.01 * c)
                i += 1
    mat = graphyt.MaterialModels()
    mat.setModel(1)
    mat.setBulkMod(1e6)
    mat.setShearMod(5e5)
    mat.setDensity(1000.0)
    prm = graphyt.Parameters(1.0, True)
    prm.setGravity(0.0, 0.0, -9.81)
    solver = graphyt.Solver(nodes, mp, prm, [mat])
    return nodes, mp, prm, solver


def test_vtp_writes_the_host_module_arrays(tmp_path):
    nodes, mp, _, _ = small_setup()
    n = mp.numParticles
    for p in range(n):
        mp.setVel(p, 0.1 * p, -0.2, 0.3)
        mp.setStress(p, 1.0 * p, 2.0, 3.0, 4.0, 5.0, 6.0)
    w = vtk.VTPWriter(n)
    w.AddPositions("Position", mp.getPosArr(), 3, vtk.Order.ijk)
    w.AddArray("Velocity", mp.getVelArr(), 3, vtk.Order.ijk)
    w.AddArray("Stress", mp.getStressArr(), 6, vtk.Order.ijk)
    w.AddArray("Material", mp.getMatArr(), 1, vtk.Order.ijk)
    f = str(tmp_path / "out" / "p.vtp")
    w.Write(f)
    r = vtk.read(f)
    assert np.array_equal(r["Position"], mp.getPosArr())
    assert np.array_equal(r["Velocity"], mp.getVelArr())
    assert np.array_equal(r["Stress"], mp.getStressArr())
    assert r["Material"].dtype == np.int32 and np.array_equal(r["Material"], mp.getMatArr())
    assert np.array_equal(r["connectivity"], np.arange(n)) and np.array_equal(r["offsets"], np.arange(1, n + 1))
    # arrays are read when Write runs: later changes to the host arrays are written
    mp.setVel(0, 9.0, 9.0, 9.0)
    w.Write(f)
    assert np.array_equal(vtk.read(f)["Velocity"][:3], [9.0, 9.0, 9.0])
    text = open(f, "rb").read().split(b"<AppendedData")[0].decode()
    assert '<VTKFile type="PolyData"' in text and 'NumberOfPoints="64"' in text


def test_vti_writes_node_fields(tmp_path):
    nodes, _, _, _ = small_setup()
    dims = [nodes.Nx, nodes.Ny, nodes.Nz]
    vel = np.arange(3 * nodes.numNodes, dtype=np.float64)
    w = vtk.VTIWriter(dims, [-0.005, -0.005, -0.005], nodes.cellsize)
    w.AddArray("Mass", nodes.getMassArr(), 1, vtk.Order.ijk)
    w.AddArray("Velocity", vel, 3, vtk.Order.ijk)
    f = str(tmp_path / "g.vti")
    w.Write(f)
    r = vtk.read(f)
    assert r["_extent"] == [0, dims[0] - 1, 0, dims[1] - 1, 0, dims[2] - 1]
    assert r["_origin"] == [-0.005] * 3 and r["_spacing"] == [0.02] * 3
    assert np.array_equal(r["Velocity"], vel) and r["Mass"].size == nodes.numNodes


def test_writer_rejects_wrong_sizes():
    w = vtk.VTPWriter(4)
    with pytest.raises(ValueError):
        w.AddArray("x", np.zeros(7), 2)
    with pytest.raises(ValueError):
        w.AddPositions("p", np.zeros(13), 3)


def test_checkpoint_format_roundtrip(tmp_path):
    arrays = {"pos": np.random.default_rng(1).random(30), "material": np.arange(10, dtype=np.int32),
              "hold": np.array([0, 1, 1], np.uint8)}
    f = str(tmp_path / "c.ckpt")
    checkpoint.write(f, arrays, t=1.25, step=17, dt=3e-6)
    back, t, step, dt = checkpoint.read(f)
    assert (t, step, dt) == (1.25, 17, 3e-6)
    for k, v in arrays.items():
        assert back[k].dtype == v.dtype and np.array_equal(back[k], v)
    with open(f, "r+b") as fh:
        fh.write(b"NOTCKPT!")
    with pytest.raises(ValueError):
        checkpoint.read(f)


def test_solver_checkpoint_and_restore(tmp_path):
    nodes, mp, _, solver = small_setup()
    n = mp.numParticles
    for p in range(n):
        mp.setVel(p, 0.01 * p, 0.02, -0.03)
        mp.setStress(p, float(p), 1.0, 2.0, 3.0, 4.0, 5.0)
    mp.holdParticle(3)
    nodes.setBC(5, graphyt.BCTypes.grid_vx, [0.25])
    f = str(tmp_path / "s.ckpt")
    solver.checkpoint(f, 0.5, 42)
    saved = {k: getattr(mp, g)().copy() for k, g in (("vel", "getVelArr"), ("stress", "getStressArr"),
                                                      ("pos", "getPosArr"), ("mass", "getMassArr"))}
    # the file is readable by the Python reader, with the reference's array names
    arrays, t, step, dt = checkpoint.read(f)
    assert (t, step) == (0.5, 42) and dt == solver.dt
    assert np.array_equal(arrays["vel"], saved["vel"]) and arrays["hold"][3] == 1
    assert arrays["node_bc_type"][5] == graphyt.BCTypes.grid_vx and arrays["node_bc_val"][20] == 0.25
    # change the state, restore, compare
    for p in range(n):
        mp.setVel(p, 1.0, 1.0, 1.0)
        mp.setStress(p, *([0.0] * 6))
    t, step = solver.restore(f)
    assert (t, step) == (0.5, 42)
    for k, g in (("vel", "getVelArr"), ("stress", "getStressArr"), ("pos", "getPosArr"), ("mass", "getMassArr")):
        assert np.array_equal(getattr(mp, g)(), saved[k]), k


def test_restore_rejects_another_setup(tmp_path):
    _, _, _, solver = small_setup(4)
    f = str(tmp_path / "s.ckpt")
    solver.checkpoint(f, 0.0, 0)
    _, _, _, other = small_setup(3)
    with pytest.raises(RuntimeError):
        other.restore(f)
