"""Learned Drucker-Prager return map: data from the GPU step, training, gated evaluation.

The network (graphyt_amd.runtime.Mlp, kernels in csrc/gy_mlp.hip, the step's evaluation
in csrc/gy_nn.hpp) gives the multiplier of the cone return for a yielding constitutive
sample; the new stress and accumulated plastic strain are rebuilt from it along the exact
return direction, so the 7 outputs of the map (new stress, accumulated plastic strain)
come from one learned scalar.

  * sample: strain increment (6), prior stress (6), accumulated plastic strain (1), as the
    plain method generates them: two consecutive states of a GPU run of the Drucker-Prager
    column, at friction angles across the trained range;
  * network input: the sample and the material's friction angle (degrees), normalised;
  * label: the exact device multiplier dL (the closed-form map, the same code the plain
    step runs) relative to the linear estimate F / (G + K eta eta_bar) of the trial yield
    function F; the network output is mapped to it by a scale and an offset.

Learning the 7 outputs directly (an earlier form of this module) left a 7e-3 relative
stress error: the plastic correction is a small difference of the large prior stress and
strain-increment terms, below what a BF16 network resolves. The multiplier relative to
its linear estimate is an O(1) quantity; for a cone without hardening (c = 0, the column)
it is 1 up to rounding, so the network's task there is to reproduce a constant over the
input range and its error is small.

Only yielding samples are learned: the gate computes the elastic trial exactly, sends
yielding samples inside the trained input range through the network and the others
through the closed form, so elastic samples are always exact.
"""
from __future__ import annotations

import math

import numpy as np

from . import runtime, workloads as W

NIN = 14          # network inputs: the 13 sample values and the friction angle
NSAMPLE = 13      # constitutive sample width
FLOP_PER_SAMPLE = 2 * (16 * 64 + 2 * 64 * 64 + 64 * 16)   # products the matrix cores run (padded)


def dp_constants(mat):
    """eta, xi, eta_bar of ConstitutiveModels.cpp:246-295 (pi = 3.14159 as there)."""
    phi = mat.phi_deg * 3.14159 / 180.0
    psi = mat.psi_deg * 3.14159 / 180.0
    eta = 6.0 * math.sin(phi) / (math.sqrt(3.0) * (3.0 - math.sin(phi)))
    xi = 6.0 * math.cos(phi) / (math.sqrt(3.0) * (3.0 - math.sin(phi)))
    etab = eta if phi == psi else 6.0 * math.sin(psi) / (math.sqrt(3.0) * (3.0 - math.sin(psi)))
    return eta, xi, etab


def trial(mat, x):
    """Elastic trial stress, yield flag and trial yield function F of raw samples (FP32
    arithmetic, as the device computes them)."""
    tr, _, _, F, yld = _trial(mat, x)
    return tr, yld, F


def _trial(mat, x):
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
    J2 = np.float32(0.5) * ((d ** 2).sum(axis=1) + np.float32(2.0) * (tr[:, 3:] ** 2).sum(axis=1))
    eta, xi, _ = dp_constants(mat)
    sJ2 = np.sqrt(J2)
    F = sJ2 + np.float32(eta) * I1 - np.float32(xi) * np.float32(mat.cohesion)
    yld = (F > 0) & (I1 < 0) if mat.is_damage else (F > 0)
    return tr, I1, sJ2, F.astype(np.float32), yld


def linear_multiplier(mat, F):
    """F / (G + K eta eta_bar): the cone multiplier without hardening."""
    eta, _, etab = dp_constants(mat)
    return np.asarray(F, np.float32) / np.float32(mat.shear_mod + mat.bulk_mod * eta * etab)


def features(mat, x):
    """Network inputs [n, 14]: the samples and the friction angle."""
    x = np.asarray(x, np.float32)
    return np.concatenate([x, np.full((x.shape[0], 1), mat.phi_deg, np.float32)], axis=1)


def collect(scene, pairs: int, gap: int, start: int = 20):
    """Constitutive samples [n, 13] from a GPU run of `scene` (the plain method): after
    `start` steps, then every `gap` steps, one pair of consecutive states (each particle
    gives a sample)."""
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


def collect_columns(make_scene, phis, pairs: int, gap: int, start: int = 20):
    """Samples of the Drucker-Prager column at each friction angle in `phis`:
    a list of (material, samples)."""
    data = []
    for phi in phis:
        sc = make_scene(phi)
        data.append((sc.mats[0], collect(sc, pairs, gap, start)))
    return data


def labels(mat, x, min_overshoot=1e-3, min_sj2=1e-15):
    """Training mask and network targets of raw samples: the exact device multiplier over
    its linear estimate, for yielding samples whose trial overshoot F is at least
    `min_overshoot` sqrt(J2). Below that the FP32 multiplier's own rounding (the
    consistency residual is a difference of terms of size sqrt(J2)) is a visible part of
    the ratio, noise the network cannot and need not learn: such samples move the stress
    by F, and the network's error on them by a fraction of that. sqrt(J2) below `min_sj2`
    (Pa) is left out too: there J2 is subnormal in FP32, flushed to zero on the device and
    not in numpy, so the two estimates of F disagree."""
    ex = runtime.dp_map(mat, x)
    _, _, sJ2, F, yld = _trial(mat, x)
    keep = yld & (sJ2 > min_sj2) & (F >= min_overshoot * sJ2)
    lin = linear_multiplier(mat, F)
    y = np.zeros(len(x), np.float32)
    y[keep] = ex["multiplier"][keep] / lin[keep]
    return keep, y


class Surrogate:
    """A trained network with its normalisation, for Drucker-Prager materials of the
    family it was trained on (friction angle an input)."""

    def __init__(self, seed: int = 0):
        self.net = runtime.Mlp(seed)
        self.stats = {}

    def fit(self, data, epochs=20, batch=4096, lr=2e-3, lr_decay=0.85, margin=0.05, min_scale=1e-3, out_q=1e-4):
        """`data`: list of (material, raw samples). Labels with the exact device map, keeps
        the yielding samples (see `labels`), normalises and trains. The gate's input range
        is that of every yielding sample, its output range that of the central targets
        (quantiles `out_q`, 1 - `out_q`), both widened by `margin`. Returns the training
        statistics."""
        xs, ys = [], []
        total = yielding = 0
        lo = hi = None
        for mat, x in data:
            x = np.asarray(x, np.float32)
            total += len(x)
            keep, y = labels(mat, x)
            xs.append(features(mat, x[keep]))
            ys.append(y[keep])
        xs, ys = np.concatenate(xs), np.concatenate(ys)
        # the gate's output range: the central targets (quantiles q, 1 - q) widened by
        # margin and two output scales; targets outside it (rounding outliers) are dropped
        if len(ys) < 64:
            raise ValueError("too few yielding samples to train on")
        qlo, qhi = (float(v) for v in np.quantile(ys, [out_q, 1.0 - out_q]))
        wide = max(margin * (qhi - qlo), 2.0 * min_scale)
        ylo, yhi = qlo - wide, qhi + wide
        inband = (ys >= ylo) & (ys <= yhi)
        dropped = int((~inband).sum())
        xs, ys = xs[inband], ys[inband]
        if len(xs) < 64:
            raise ValueError("too few yielding samples to train on")
        mean, std = xs.mean(axis=0), xs.std(axis=0)
        std = np.where(std > 0, std, 1.0).astype(np.float32)
        xn = ((xs - mean) / std).astype(np.float32)
        for mat, x in data:   # the gate's range: every yielding sample, trained or not
            x = np.asarray(x, np.float32)
            yld = trial(mat, x)[1]
            yielding += int(yld.sum())
            f = (features(mat, x[yld]) - mean) / std
            if len(f):
                lo = f.min(axis=0) if lo is None else np.minimum(lo, f.min(axis=0))
                hi = f.max(axis=0) if hi is None else np.maximum(hi, f.max(axis=0))
        span = hi - lo
        lo, hi = lo - margin * span, hi + margin * span
        off = float(ys.mean())
        scale = float(max(ys.std(), min_scale))
        tn = ((ys - off) / scale).astype(np.float32)[:, None]
        self.net.set_norm(mean, std, lo, hi, scale, off, ylo, yhi)
        order = np.random.default_rng(0).permutation(len(xn))
        loss = self.net.train(xn[order], tn[order], epochs=epochs, batch=batch, lr=lr, lr_decay=lr_decay)
        self.stats = {"samples": int(total), "yielding": yielding, "trained": int(len(xs)),
                      "dropped_outside_output_range": dropped, "final_loss": float(loss),
                      "target_mean": off, "target_std": float(ys.std()), "out_scale": scale,
                      "out_range": [ylo, yhi]}
        return self.stats

    def apply(self, mat, samples) -> dict:
        return runtime.dp_map(mat, samples, self.net)

    def close(self):
        self.net.close()


def rel_l2(a, b) -> float:
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def evaluate(sur: Surrogate, mat, samples, reps: int = 5) -> dict:
    """The gated network against the closed-form device map on the same samples: relative
    errors (stress over all samples and the yielding ones, accumulated plastic strain and
    multiplier over the yielding ones), how many took the network, and kernel times (best
    of `reps`) of the closed form, the reference's loop and the network, with the
    network's matrix-core rate."""
    x = np.asarray(samples, np.float32)
    ex = [runtime.dp_map(mat, x) for _ in range(reps)]
    lp = [runtime.dp_map(mat, x, mode=runtime.DP_REFERENCE) for _ in range(reps)]
    su = [sur.apply(mat, x) for _ in range(reps)]
    e, s = ex[0]["out"], su[0]["out"]
    _, yld, _ = trial(mat, x)
    t_exact = min(r["ms"] for r in ex)
    t_loop = min(r["ms"] for r in lp)
    t_surr = min(r["ms"] for r in su)
    return {
        "samples": int(len(x)), "yielding": int(yld.sum()),
        "network": su[0]["network"], "fallback": su[0]["fallback"],
        "stress_rel_l2_all": rel_l2(s[:, :6], e[:, :6]),
        "stress_rel_l2_yielding": rel_l2(s[yld, :6], e[yld, :6]) if yld.any() else 0.0,
        "epbar_rel_l2_yielding": rel_l2(s[yld, 6], e[yld, 6]) if yld.any() else 0.0,
        "multiplier_rel_l2": rel_l2(su[0]["multiplier"][yld], ex[0]["multiplier"][yld]) if yld.any() else 0.0,
        "loop_vs_closed_stress_rel_l2": rel_l2(lp[0]["out"][:, :6], e[:, :6]),
        "closed_ms": t_exact, "reference_loop_ms": t_loop, "network_ms": t_surr,
        "speedup_vs_closed": t_exact / t_surr, "speedup_vs_reference_loop": t_loop / t_surr,
        "network_tflops": su[0]["network"] * FLOP_PER_SAMPLE / (t_surr * 1e-3) / 1e12,
    }


def column_material(phi=30.0, dp_return=0):
    """The Drucker-Prager column's material (model 4, rho 2000, K 5e7, G 2e7, c 0, psi 0,
    guarded apex) at friction angle phi."""
    return W.material(model=4, bulk_mod=5e7, shear_mod=2e7, density=2000.0, cohesion=0.0, phi=phi, psi=0.0,
                      critical_strain=0.01, dp_apex_guard=True, dp_return=dp_return)
