"""The graphyt host module on a plane setup (Parameters(tmax, False), L[2] = 0), without
GPU calls: one node layer (Nodes.cpp:71-73), and Solver::initialize's per-unit-depth mass
and volume with no unit-mass override for rigid materials (Solver.cpp:57-66)."""
import pytest

from graphyt_amd import graphyt


def test_plane_grid_and_initialize():
    nodes = graphyt.Nodes([0.2, 0.3, 0.0], 0.02)
    assert (nodes.Nx, nodes.Ny, nodes.Nz) == (11, 16, 1) and nodes.numNodes == 11 * 16
    mp = graphyt.MaterialPoints(nodes)
    mp.psep = 0.01
    for i in range(4):
        mp.add(i, 0.05 + 0.01 * i, 0.05, 0.0)
    mp.setMaterialID(3, 1)
    rock = graphyt.MaterialModels()
    rock.setDensity(2600.0)
    platen = graphyt.MaterialModels()
    platen.setIsRigid(True)
    prm = graphyt.Parameters(1.0, False)
    graphyt.Solver(nodes, mp, prm, [rock, platen])
    mass, vol = mp.getMassArr(), mp.getVolumeArr()
    assert mass[0] == pytest.approx(2600.0 * 1e-4) and vol[0] == pytest.approx(1e-4)
    assert mass[3] == pytest.approx(1000.0 * 1e-4)  # rigid keeps rho * V in the plane path


def test_pyck_int_fields_set_material_and_object():
    """addPyckIntField (MaterialPoints.cpp:32-53): "material" and "objectID" fields set the
    particle arrays, other names are ignored; addPyckDoubleField does nothing."""
    nodes = graphyt.Nodes([0.2, 0.2, 0.2], 0.02)
    mp = graphyt.MaterialPoints(nodes)
    mp.psep = 0.01
    mp.addPyckGeom([0.05, 0.05, 0.05, 0.06, 0.05, 0.05, 0.07, 0.05, 0.05], 3)
    mat = graphyt.IntField("material", 3, 1)
    mat.data[:] = [0, 1, 1]
    obj = graphyt.IntField("objectID", 3)
    obj.data[:] = [4, 4, 7]
    other = graphyt.IntField("color", 3)
    other.data[:] = 9
    for f in (mat, obj, other):
        mp.addPyckIntField(f)
    mp.addPyckDoubleField(None)
    assert list(mp.getMatArr()) == [0, 1, 1]
    assert [mp.getObjID(i) for i in range(3)] == [4, 4, 7]
