"""Full-size parity run of the Drucker-Prager column (1,024,000 particles, h = 0.01,
guarded apex): the GPU step against the FP64 oracle restatement on the host's cores, after
`steps` steps, relative L2 over all particles per field. Not collected by pytest (the
oracle takes ~0.4 s per step at this size); run as

    python3 tests/full_parity_column.py [steps [result.json]]
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from graphyt_amd import runtime, workloads as W  # noqa: E402

FIELDS = ["pos", "vel", "stress", "density", "acc_plastic_strain", "strain"]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def main(steps: int = 1000, out_path: str = "") -> int:
    sc = W.dp_column()
    t0 = time.time()
    g = runtime.Context(sc)
    g.run(steps)
    a = g.download(FIELDS)
    print(f"gpu: {steps} steps in {time.time() - t0:.2f} s (incl. upload)", flush=True)
    o = oracle.CpuSim(sc, "restatement")
    for s0 in range(0, steps, 100):
        n = min(100, steps - s0)
        o.run(n, step0=s0)
        print(f"oracle: {s0 + n} steps, {time.time() - t0:.0f} s", flush=True)
    b = o.get(FIELDS)
    out = {k: rel(a[k], b[k]) for k in FIELDS}
    out.update(steps=steps, particles=sc.n, max_abs_pos_diff=float(np.abs(a["pos"] - b["pos"]).max()),
               nonfinite_gpu=int(np.isnan(a["stress"]).sum()), nonfinite_oracle=int(np.isnan(b["stress"]).sum()))
    print(json.dumps(out), flush=True)
    if out_path:
        os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
        with open(out_path, "w") as f:
            json.dump(out, f)
    return 0


if __name__ == "__main__":
    sys.exit(main(int(sys.argv[1]) if len(sys.argv) > 1 else 1000, sys.argv[2] if len(sys.argv) > 2 else ""))
