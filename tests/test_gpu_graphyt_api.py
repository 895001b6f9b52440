"""The graphyt surface on the GPU: a reference-style script (Nodes, MaterialPoints.add,
MaterialModels, Parameters, Solver.iterate_GPU in a Python step loop) against the FP64
oracle on the same scene, lazy host/device synchronisation, node fields and the
graph-replayed step loop. Relative L2 tolerances as in test_gpu_parity."""
import numpy as np
import pytest

from graphyt_amd import graphyt, workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def build_cube(sc):
    """The elastic-cube configuration written the way the reference's scripts are."""
    nodes = graphyt.Nodes(list(sc.L), sc.h)
    mp = graphyt.MaterialPoints(nodes)
    mp.psep = sc.psep
    pos, vel = sc.particles["pos"], sc.particles["vel"]
    for i in range(sc.n):
        mp.add(i, *pos[i])
    for i in range(sc.n):
        mp.setVel(i, *vel[i])
    m = graphyt.MaterialModels()
    m.setBulkMod(1e6)
    m.setShearMod(5e5)
    m.setDensity(1000.0)
    prm = graphyt.Parameters(1.0, True)
    prm.setGravity(*sc.gravity)
    solver = graphyt.Solver(nodes, mp, prm, [m])
    solver.setDt(sc.resolved_dt())
    return nodes, mp, prm, solver


def test_reference_style_step_loop_matches_oracle(oracle_mod):
    sc = W.elastic_cube(steps=200)
    nodes, mp, prm, solver = build_cube(sc)
    for step in range(200):
        solver.iterate_GPU(step * solver.dt, step, prm)
    o = oracle_mod.CpuSim(sc, "restatement")
    o.run(200)
    b = o.get(["pos", "vel", "stress", "density"])
    assert rel(mp.getPosArr().reshape(-1, 3), b["pos"]) <= 1e-6
    assert rel(mp.getVelArr().reshape(-1, 3), b["vel"]) <= 1e-4
    assert rel(mp.getStressArr().reshape(-1, 6), b["stress"]) <= 1e-4
    assert rel(mp.getDensityArr(), b["density"]) <= 1e-6
    # displacement = x - x0 in the caller's particle order
    assert np.allclose(mp.getDispArr().reshape(-1, 3), mp.getPosArr().reshape(-1, 3) - sc.particles["pos"],
                       atol=1e-12)


def test_setter_between_steps_and_node_fields():
    sc = W.elastic_cube(steps=20)
    nodes, mp, prm, solver = build_cube(sc)
    for step in range(10):
        solver.iterate_GPU(step * solver.dt, step, prm)
    v_before = mp.getVelArr().reshape(-1, 3).copy()
    mp.setVel(0, 1.0, 2.0, 3.0)                     # pulls the device state, then edits it
    assert np.array_equal(mp.getVelArr().reshape(-1, 3)[1:], v_before[1:])
    solver.iterate_GPU(10 * solver.dt, 10, prm)      # uploads the edited state
    assert mp.getVelX(0) != v_before[0, 0]
    # nodal mass of the last step sums to the particle mass (periodic classes counted once
    # in the canonical nodes: the mirror nodes hold copies)
    m_nodes = nodes.getMassArr()
    nx = nodes.Nx
    canon = np.zeros((nodes.Nz, nodes.Ny, nx), bool)
    canon[:nodes.Nz - 3, :nodes.Ny - 3, :nx - 3] = True
    assert m_nodes[canon.ravel()].sum() == pytest.approx(mp.getMassArr().sum(), rel=1e-5)
    f = nodes.getForceArr().reshape(-1, 3)
    assert np.isfinite(f).all()


def test_graph_replayed_steps_match_python_loop():
    sc = W.elastic_cube(steps=40)
    _, mp1, prm1, s1 = build_cube(sc)
    for step in range(40):
        s1.iterate_GPU(step * s1.dt, step, prm1)
    _, mp2, prm2, s2 = build_cube(sc)
    s2.iterate_GPU_steps(40, prm2)
    assert rel(mp2.getPosArr(), mp1.getPosArr()) <= 1e-7
    assert rel(mp2.getStressArr(), mp1.getStressArr()) <= 1e-4


def test_plane_dam_break_script_matches_reference_fixture():
    """Example_Scripts/dambreak_2D.py written against the graphyt module (Parameters(tmax,
    False), L[2] = 0, wall nodes held by grid_vx_vy), stepped with iterate_GPU: the plane
    path, compared with the fixture the reference wrote for the same scene."""
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import scenes
    sc, steps, _ = scenes.build("plane_dam_break")
    nodes = graphyt.Nodes(list(sc.L), sc.h)
    assert nodes.Nz == 1 and nodes.numNodes == sc.num_nodes
    mp = graphyt.MaterialPoints(nodes)
    mp.psep = sc.psep
    for i, p in enumerate(sc.particles["pos"]):
        mp.add(i, *p)
    water = graphyt.MaterialModels()
    water.setBulkMod(1e6)
    water.setShearMod(0.0)
    water.setDensity(1000)
    water.setModel(2)
    wall = 2.0 * nodes.cellsize
    for n in range(nodes.numNodes):
        x, y = nodes.getPosX(n), nodes.getPosY(n)
        if x >= sc.L[0] - wall or x <= wall or y >= sc.L[1] - wall or y <= wall:
            nodes.setBC(n, graphyt.BCTypes.grid_vx_vy, [0, 0])
    prm = graphyt.Parameters(5.0, False)
    prm.setGravity(0, -10, 0)
    prm.setDampingCoef(0.5)
    solver = graphyt.Solver(nodes, mp, prm, [water])
    solver.setDt(sc.resolved_dt())
    assert mp.getMassArr()[0] == pytest.approx(1000.0 * sc.psep ** 2)  # per unit depth
    for step in range(steps):
        solver.iterate_GPU(step * solver.dt, step, prm)
    with np.load(os.path.join(here, "golden", "plane_dam_break.npz")) as f:
        fx = {k: f[k] for k in f.files}
    assert rel(mp.getPosArr().reshape(-1, 3), fx["pos"]) <= 1e-6
    assert rel(mp.getVelArr().reshape(-1, 3), fx["vel"]) <= 1e-4
    assert rel(mp.getStressArr().reshape(-1, 6), fx["stress"]) <= 1e-4
