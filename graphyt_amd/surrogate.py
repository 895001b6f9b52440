"""Learned Drucker-Prager return map: data from the GPU step, training, gated evaluation.

The network (graphyt_amd.runtime.Mlp, kernels in csrc/gy_mlp.hip) maps a constitutive
sample -- strain increment (6), prior stress (6), accumulated plastic strain (1) -- to the
new stress (6) and accumulated plastic strain (1). It is trained on stress updates the
plain method generates:

  * inputs: the states of Drucker-Prager column runs on the GPU, two consecutive steps
    apart (strain increment = difference of the total strains, prior stress and
    accumulated plastic strain of the first);
  * labels: the exact FP32 return map (the same device code the step runs) on those
    inputs.

Only yielding samples (positive trial yield function) are learned: the gate computes the
elastic trial exactly, sends yielding samples inside the trained input range through the
network and the others through the exact map, so elastic samples are always exact.
"""
from __future__ import annotations

import numpy as np

from . import runtime, workloads as W

NIN, NOUT = 13, 7


def trial(mat, x):
    """Elastic trial stress and yield flag of raw samples (FP32, as the device computes)."""
    x = np.asarray(x, np.float32)
    G, K = np.float32(mat.shear_mod), np.float32(mat.bulk_mod)
    de, s = x[:, 0:6], x[:, 6:12]
    vol = np.float32(0.3333) * (de[:, 0] + de[:, 1] + de[:, 2])
    tr = np.empty_like(s)
    for k in range(3):
        tr[:, k] = s[:, k] + (np.float32(2.0) * G * (de[:, k] - vol) + np.float32(3.0) * K * vol)
    for k in range(3, 6):
        tr[:, k] = s[:, k] + np.float32(2.0) * G * de[:, k]
    I1 = (tr[:, 0] + tr[:, 1] + tr[:, 2]) / np.float32(3.0)
    d = tr[:, :3] - I1[:, None]
    J2 = 0.5 * ((d ** 2).sum(axis=1) + 2.0 * (tr[:, 3:] ** 2).sum(axis=1))
    phi = mat.phi_deg * 3.14159 / 180.0  # pi as ConstitutiveModels.cpp:247-248
    eta = 6.0 * np.sin(phi) / (np.sqrt(3.0) * (3.0 - np.sin(phi)))
    xi = 6.0 * np.cos(phi) / (np.sqrt(3.0) * (3.0 - np.sin(phi)))
    F = np.sqrt(J2) + eta * I1 - xi * mat.cohesion
    yld = (F > 0) & (I1 < 0) if mat.is_damage else (F > 0)
    return tr, yld


def collect(scene, pairs: int, gap: int, start: int = 20):
    """Constitutive samples [n, 13] from a GPU run of `scene`: after `start` steps, then
    every `gap` steps, one pair of consecutive states (each particle gives a sample)."""
    ctx = runtime.Context(scene)
    ctx.run(start)
    out, step = [], start
    fields = ["strain", "stress", "acc_plastic_strain"]
    for _ in range(pairs):
        a = ctx.download(fields)
        ctx.run(1, step0=step)
        step += 1
        b = ctx.download(fields)
        x = np.concatenate([b["strain"] - a["strain"], a["stress"], a["acc_plastic_strain"][:, None]], axis=1)
        out.append(x.astype(np.float32))
        ctx.run(gap, step0=step)
        step += gap
    ctx.close()
    return np.concatenate(out)


class Surrogate:
    """A trained network plus its normalisation, for one material."""

    def __init__(self, mat, seed: int = 0):
        self.mat = mat
        self.net = runtime.Mlp(seed)
        self.stats = {}

    def fit(self, samples, epochs=12, batch=4096, lr=2e-3, lr_decay=0.8, margin=0.05):
        """Label `samples` with the exact device map, keep the yielding ones, normalise and
        train. Returns the training statistics."""
        x = np.asarray(samples, np.float32)
        exact = runtime.dp_map(self.mat, x)["out"]
        tr, yld = trial(self.mat, x)
        xs, ys, trs = x[yld], exact[yld], tr[yld]
        if len(xs) < 64:
            raise ValueError("too few yielding samples to train on")
        mean, std = xs.mean(axis=0), xs.std(axis=0)
        std = np.where(std > 0, std, 1.0).astype(np.float32)
        xn = (xs - mean) / std
        lo, hi = xn.min(axis=0), xn.max(axis=0)
        span = hi - lo
        lo, hi = lo - margin * span, hi + margin * span
        resid = np.concatenate([ys[:, :6] - trs, (ys[:, 6] - xs[:, 12])[:, None]], axis=1)
        scale = resid.std(axis=0)
        scale = np.where(scale > 0, scale, 1.0).astype(np.float32)
        tn = resid / scale
        self.net.set_norm(mean, std, lo, hi, scale)
        rng = np.random.default_rng(0)
        order = rng.permutation(len(xn))
        loss = self.net.train(xn[order], tn[order], epochs=epochs, batch=batch, lr=lr, lr_decay=lr_decay)
        self.stats = {"samples": int(len(x)), "yielding": int(len(xs)), "final_loss": float(loss)}
        return self.stats

    def apply(self, samples) -> dict:
        return runtime.dp_map(self.mat, samples, self.net)

    def close(self):
        self.net.close()


def rel_l2(a, b) -> float:
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def evaluate(sur: Surrogate, samples, reps: int = 5) -> dict:
    """Surrogate against the exact device map on the same samples: relative stress error
    (all samples, and the yielding ones), how many took the network, and kernel times
    (best of `reps`)."""
    x = np.asarray(samples, np.float32)
    ex = [runtime.dp_map(sur.mat, x) for _ in range(reps)]
    su = [sur.apply(x) for _ in range(reps)]
    e, s = ex[0]["out"], su[0]["out"]
    _, yld = trial(sur.mat, x)
    t_exact = min(r["ms"] for r in ex)
    t_surr = min(r["ms"] for r in su)
    return {
        "samples": int(len(x)), "yielding": int(yld.sum()),
        "network": su[0]["network"], "fallback": su[0]["fallback"],
        "stress_rel_l2_all": rel_l2(s[:, :6], e[:, :6]),
        "stress_rel_l2_yielding": rel_l2(s[yld, :6], e[yld, :6]) if yld.any() else 0.0,
        "epbar_rel_l2_yielding": rel_l2(s[yld, 6], e[yld, 6]) if yld.any() else 0.0,
        "exact_ms": t_exact, "surrogate_ms": t_surr, "speedup": t_exact / t_surr,
    }


def column_material(phi=30.0):
    """The Drucker-Prager column's material (model 4, rho 2000, K 5e7, G 2e7, c 0, psi 0,
    guarded apex) at friction angle phi."""
    return W.material(model=4, bulk_mod=5e7, shear_mod=2e7, density=2000.0, cohesion=0.0, phi=phi, psi=0.0,
                      critical_strain=0.01, dp_apex_guard=True)
