"""Serialise after GPU steps (MaterialPoints.cpp:341-364): the host module's CSV after 40
iterate_GPU steps against the CSV the reference's own Serialise writes for the reference's
Solver stepped through its iterate_GPU slot (then Synchronise): same header, rows, order
and number format, values equal to the GPU step's run-to-run reproducibility (the order
of atomic adds can move a node sum by one FP32 ulp; on this cube two runs differ by
~1e-5 of a column's largest value at most); and, rows matched by position (the CPU run has
sorted its particles by cell), the reference's own 40 iterate_CPU steps within 2e-5 of
each column's largest value."""
import re
import os

import numpy as np
import pytest

from test_serialise import host_module_setup, read, scene

pytestmark = pytest.mark.gpu


def rows(path):
    a = np.loadtxt(path, delimiter=",", skiprows=1)
    return a[np.lexsort((a[:, 2], a[:, 1], a[:, 0]))]


def test_serialise_after_gpu_steps(oracle_mod, tmp_path):
    if not oracle_mod.available("reference_gpu"):
        pytest.skip("oracle/_ref/libgraphyt_ref_gpu.so not built (needs the reference sources)")
    sc = scene()
    cwd = os.getcwd()
    try:
        os.chdir(tmp_path)
        os.makedirs("output")
        nodes, mp, prm, solver = host_module_setup(sc)
        for step in range(40):
            solver.iterate_GPU(step * solver.dt, step, prm)
        mp.Serialise(40, "ours")
        slot = oracle_mod.CpuSim(sc, "reference_gpu")
        slot.run_gpu(40)
        slot.synchronise()
        slot.serialise(40, "slot")
        slot.close()
        cpu = oracle_mod.CpuSim(sc, "reference_gpu")
        cpu.run(40)
        cpu.serialise(40, "cpu")
        cpu.close()
    finally:
        os.chdir(cwd)
    out = tmp_path / "output"
    ours, theirs = read(out / "ours_mpts_40.csv").splitlines(), read(out / "slot_mpts_40.csv").splitlines()
    assert ours[0] == theirs[0] and len(ours) == len(theirs) == sc.n + 1
    fmt = re.compile(r"^-?\d+\.\d{6}(,-?\d+\.\d{6}){9}$")  # ten %f fields
    assert all(fmt.match(line) for line in ours[1:]) and all(fmt.match(line) for line in theirs[1:])
    a_, b_ = (np.loadtxt(out / f, delimiter=",", skiprows=1) for f in ("ours_mpts_40.csv", "slot_mpts_40.csv"))
    assert np.allclose(a_[:, :3], b_[:, :3], rtol=0, atol=2e-6)  # same particles in the same order
    rep = np.abs(a_ - b_).max(axis=0) / np.maximum(np.abs(b_).max(axis=0), 1e-30)
    print("run-to-run, max abs difference / column max:", np.array2string(rep, precision=2), flush=True)
    assert (rep <= 2e-5).all()
    a, b = rows(out / "ours_mpts_40.csv"), rows(out / "cpu_mpts_40.csv")
    assert a.shape == b.shape == (sc.n, 10)
    assert not np.allclose(a[:, 6:9], 0.0)  # the cube moved
    err = np.abs(a - b).max(axis=0) / np.maximum(np.abs(b).max(axis=0), 1e-30)
    print("max abs difference / column max:", np.array2string(err, precision=2), flush=True)
    assert (err <= 2e-5).all()
