"""Write the golden fixtures of this folder from the reference's own code.

The reference's C++ sources (Solver, Grid, Extrapolation, TimeIntegration,
ConstitutiveModels, DamageModels, eig3, ...) are compiled in place by oracle/Makefile
into oracle/_ref/libgraphyt_ref.so, with a small harness (oracle/ref_harness.cpp) that
builds the scene through Nodes, MaterialPoints, Solver::setDt and steps it with
Solver::iterate_CPU. For every scene of scenes.py this script runs that library and
stores, per particle in the caller's order, the FP64 fields after the scene's steps:

    tests/golden/<name>.npz: the fields of scenes.fields(name), "steps", "digest"

Run in a checkout that has the reference sources (python tests/golden/make_golden.py);
the fixtures are then committed and read anywhere without them.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle  # noqa: E402
import scenes  # noqa: E402


def run(sim, steps, t_of):
    if t_of is None:
        sim.run(steps)
    else:
        for s in range(steps):
            sim.iterate(t_of(s), s)


def main(names=None):
    oracle.build(reference=True)
    if not oracle.available("reference"):
        raise SystemExit("the reference library could not be built (no reference sources here)")
    for name in names or scenes.ALL_NAMES:
        sc, steps, t_of = scenes.build(name)
        sim = oracle.CpuSim(sc, "reference")
        run(sim, steps, t_of)
        out = sim.get(scenes.fields(name))
        sim.close()
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, steps=np.int64(steps), digest=scenes.input_digest(sc),
                            **{k: np.asarray(v) for k, v in out.items()})
        print(f"{name}: {sc.n} particles, {steps} steps -> {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main(sys.argv[1:] or None)
