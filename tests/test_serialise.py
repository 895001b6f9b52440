"""Serialise (MaterialPoints.cpp:341-364, Nodes.cpp:165-177): the graphyt host module's CSV
writers against the files the compiled reference itself writes, for the same scene set up
through each side's Nodes / MaterialPoints / MaterialModels / Parameters / Solver.
Byte-identical at set-up (particles in the caller's order: the reference has not sorted
yet). Needs the reference build (oracle/_ref), made where the reference sources exist."""
import os

import numpy as np
import pytest

from graphyt_amd import graphyt, workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


def scene():
    """The 20 x 20 x 20-particle elastic cube with a smooth stress field."""
    sc = W.elastic_cube(steps=40)
    x = sc.particles["pos"] - 0.25
    st = np.zeros((sc.n, 6))
    st[:, 0], st[:, 1], st[:, 2] = 1e4 * x[:, 0], -2e4 * x[:, 1], 5e3 * x[:, 2]
    st[:, 3], st[:, 4], st[:, 5] = 3e3 * x[:, 1], 1e3 * x[:, 2], -2e3 * x[:, 0]
    sc.particles.arrays["stress"] = st
    # zero: no heat transfer. With a temperature the reference's own run would scramble the
    # particles' heat-content gradients at its sorts (a documented deviation this step does
    # not reproduce), and the T column would differ for that reason alone.
    sc.particles.arrays["temperature"] = np.zeros(sc.n)
    return sc


def host_module_setup(sc):
    L = list(sc.L)
    nodes = graphyt.Nodes(L, sc.h)
    mp = graphyt.MaterialPoints(nodes)
    mp.psep = sc.psep
    pos, vel, st, T = (sc.particles[k] for k in ("pos", "vel", "stress", "temperature"))
    for i in range(sc.n):
        mp.add(i, *pos[i])
    for i in range(sc.n):
        mp.setVel(i, *vel[i])
        mp.setStress(i, *st[i])
        mp.setTemp(i, T[i])
    m = sc.mats[0]
    mat = graphyt.MaterialModels()
    mat.setModel(m.model); mat.setBulkMod(m.bulk_mod); mat.setShearMod(m.shear_mod); mat.setDensity(m.density)
    mat.setViscosity(m.viscosity); mat.setYieldStress(m.yield_stress); mat.setCohesion(m.cohesion)
    mat.setCriticalStrain(m.critical_strain); mat.setFrictionAngle(m.phi_deg); mat.setDilatancyAngle(m.psi_deg)
    mat.setThermalConductivity(m.thermal_conductivity); mat.setSpecificHeatCapacity(m.specific_heat)
    mat.setTurbulentStress(m.turb_stress); mat.setIsRigid(float(m.is_rigid)); mat.setIsDamage(bool(m.is_damage))
    mat.setDamageModel(m.damage_model); mat.setCrackSpeed(m.crack_speed); mat.setDamageM(m.damage_m)
    mat.setDamageK(m.damage_k)
    prm = graphyt.Parameters(1.0, True)
    prm.setGravity(*sc.gravity)
    solver = graphyt.Solver(nodes, mp, prm, [mat])
    solver.setDt(sc.resolved_dt())
    return nodes, mp, prm, solver


def read(path):
    with open(path) as f:
        return f.read()


def test_serialise_matches_the_reference_writers(oracle_mod, tmp_path):
    if not oracle_mod.available("reference"):
        pytest.skip("the compiled reference (oracle/_ref) needs the reference sources at build time")
    sc = scene()
    cwd = os.getcwd()
    try:
        os.chdir(tmp_path)
        os.makedirs("output")
        ref = oracle_mod.CpuSim(sc, "reference")
        ref.serialise(0, "ref")
        ref.close()
        nodes, mp, prm, solver = host_module_setup(sc)
        mp.Serialise(0, "ours")
        nodes.Serialise(0, "ours")
    finally:
        os.chdir(cwd)
    out = tmp_path / "output"
    a, b = read(out / "ours_mpts_0.csv"), read(out / "ref_mpts_0.csv")
    assert a.splitlines()[0] == "x,y,z,vm,P,T,vx,vy,vz,rho"
    assert len(a.splitlines()) == sc.n + 1
    assert a == b
    assert read(out / "ours_nds_0.csv") == read(out / "ref_nds_0.csv")
