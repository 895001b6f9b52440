#!/usr/bin/env python3
"""Benchmark of the MI355X MPM step on the 3D Drucker-Prager granular column collapse.

Workload: 2x1x1 m domain, cellsize 0.01 (201x101x101 nodes), a 0.4x0.8x0.4 m column of
80x160x80 = 1,024,000 particles at psep = 0.005 with seeded U(-0.05, 0.05)*psep jitter,
model 4 (rho 2000, K 5e7, G 2e7, c 0, phi 30, psi 0, guarded apex), floor/wall nodal BCs,
gravity -9.81, dt = 0.1*h/sqrt(K/rho). Synthetic data generated in-process.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints ONE JSON line on stdout (everything else goes to stderr). `value` is whole-job
particle-steps/s with the state resident in HBM; `roofline` is the dominant kernel's
algorithmic bytes per launch over its average duration, timed with HIP events on the
context's own stream during the timed region; `cpu_baseline` times the FP64 OpenMP
restatement of the reference's iterate_CPU (its loops, schedules and sort cadence, with
the guarded apex: the reference's own map turns this column NaN) on a 60-step sample of
the same workload on this host's cores, kind "port". `surrogate` is the learned return
map at scale: the 4,194,304-particle column (h = 0.00625) stepped with the closed-form,
the reference's 201-pass loop and the learned multipliers, the network trained on 20.5M
samples of plain GPU runs at friction angles 25-40 degrees (--no-surrogate leaves it out;
--workload surrogate prints it alone).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, published
METRIC = "particle-steps/s (3D GIMP 1M pts), % HBM roofline, surrogate vs plain speedup"

# Algorithmic HBM bytes per particle-step (FP32 SoA; one compulsory read and write of every
# persistent field, binning key/offset/permutation, grid traffic amortised at ~1/8 active
# node per particle), Drucker-Prager: x 24 + v 24 + sigma 48 + eps 48 + eps_p 48 +
# eps_bar 8 + rho 8 + m 4 + material 4 + binning 12 + grid 13 = 241.
B_ALG_STEP = 241
# The fused particle kernel carries all of it except the binning offset read and
# permutation write (8 B), which the scan/place kernels carry.
B_ALG_G2P2G = B_ALG_STEP - 8


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "g2p2g_dp_column_1M.md")
PLANE_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "g2p2g_plane_dam_break_0.94M.md")


def measured_traffic(kernel: str, profile: str = TRAFFIC_PROFILE):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE passes scaled by a plain-copy calibration), or None."""
    try:
        text = open(profile).read()
        blob = text.split("```json")[-1].split("```")[0]
        rec = json.loads(blob)
        return int(rec["traffic_bytes"]) if rec.get("kernel") == kernel else None
    except (OSError, ValueError, KeyError, IndexError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(sample_steps: int, warm: int, workload: str = "dp_column") -> dict:
    """Time the CPU reference step on the same workload (run in a child process).

    dp_column: the FP64 OpenMP restatement of iterate_CPU (oracle/cpu_ref.cpp: the
    reference's loops, schedule clauses and every-10-steps sort) with the guarded apex,
    kind "port". The compiled reference cannot run this workload: with c = 0 and psi = 0
    its apex return is 0*inf (ConstitutiveModels.cpp:364-365) and the column's stresses
    turn NaN within a few steps, after which the gather does almost no work.
    plane_dam_break: the compiled reference itself (kind "reference") where it is built.
    A sample whose final state holds a non-finite value is not reported as a rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    from graphyt_amd import workloads as W
    if workload == "plane_dam_break":
        kind = "reference" if oracle.available("reference") else "restatement"
        sc = plane_scene()
    else:
        kind = "restatement"
        sc = W.dp_column()
    sim = oracle.CpuSim(sc, kind)
    for s in range(warm):
        sim.iterate(s * sim.dt, s)
    t0 = time.perf_counter()
    for s in range(warm, warm + sample_steps):
        sim.iterate(s * sim.dt, s)
    dt = time.perf_counter() - t0
    st = sim.get(["pos", "vel", "stress"])
    sim.close()
    finite = all(bool(np.isfinite(st[k]).all()) for k in st)
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    out = {"value": sc.n * sample_steps / dt if finite else None, "unit": "particle-steps/s", "cores": threads,
           "kind": "reference" if kind == "reference" else "port",
           "sample": f"{workload} {sc.n:,} particles, steps {warm}..{warm + sample_steps - 1} "
                     f"({sample_steps} steps incl. the every-10-steps sort), FP64, {dt:.1f} s"
                     + ("" if kind == "reference" else ", guarded apex (the reference's map is NaN here)"),
           "nproc": os.cpu_count(), "OMP_NUM_THREADS": threads,
           "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"), "OMP_PLACES": os.environ.get("OMP_PLACES"),
           "final_state_finite": finite}
    return out


def run_cpu_baseline(cpu_steps: int, workload: str = "dp_column"):
    """The CPU leg in a child process with the threads pinned to the host's cores
    (OMP_PROC_BIND=close, OMP_PLACES=cores unless set)."""
    env = dict(os.environ)
    env.setdefault("OMP_PROC_BIND", "close")
    env.setdefault("OMP_PLACES", "cores")
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--workload", workload,
                            "--cpu-steps", str(cpu_steps)], capture_output=True, text=True, timeout=900, env=env)
        return json.loads(r.stdout.strip().splitlines()[-1])
    except Exception as e:  # the GPU figure stands on its own
        log("cpu baseline failed:", e)
        return None


def copy_calibration(nbytes: int = 1 << 30, reps: int = 5) -> int:
    """Plain device copies of a known size for the HBM counters: a vectorised copy and a
    gather (index_select with the identity index: 4-byte loads per lane, the access width
    of the fused kernel's state loads). Each kernel's FETCH_SIZE / WRITE_SIZE per dispatch,
    divided by its bytes, is the scale applied to the counters of kernels of that access
    width."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    idx = torch.arange(src.numel(), device="cuda", dtype=torch.int64)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        torch.index_select(src, 0, idx, out=dst)
    torch.cuda.synchronize()
    log(f"copy calibration: {reps} copies and {reps} gathers of {nbytes} bytes (gather also reads "
        f"{idx.numel() * 8} index bytes)")
    print(json.dumps({"copy_bytes": nbytes, "gather_index_bytes": idx.numel() * 8, "reps": reps}))
    return 0


ENSEMBLE_STEPS = 500  # the batched-impact configuration is quoted over 500 steps


def bench_ensemble(args) -> int:
    """256 independent perfect-plastic block impacts advanced together (one context, one
    graph per step); value = total particle-steps/s over all instances. The warm-up and the
    two timed runs (events, then graph replay) stay within the configuration's 500 steps:
    later the fastest blocks, rebounding from the wall, leave their small grids."""
    from graphyt_amd import runtime, workloads as W
    scenes = W.impact_ensemble()
    n = sum(sc.n for sc in scenes)
    log(f"workload impact_ensemble: {len(scenes)} instances, {n} particles")
    if args.warmup + 2 * args.steps > ENSEMBLE_STEPS:
        args.steps = max(1, (ENSEMBLE_STEPS - args.warmup) // 2)
        log(f"timed steps per run capped at {args.steps} ({ENSEMBLE_STEPS}-step configuration)")
    ctx = runtime.BatchContext(scenes)
    tuned = ctx.autotune(10) if args.autotune else None
    if tuned:
        log("autotune:", tuned)
    ctx.run(args.warmup)
    ctx.synchronize()
    ctx.set_option(runtime.OPT_PROFILE, 1)
    ctx.reset_kernel_stats()
    t0 = time.perf_counter()
    ctx.run(args.steps, step0=args.warmup)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    stats = ctx.kernel_stats()
    ctx.set_option(runtime.OPT_PROFILE, 0)
    t0 = time.perf_counter()
    ctx.run(args.steps, step0=args.warmup + args.steps)
    ctx.synchronize()
    wall_graph = time.perf_counter() - t0
    audit = ctx.audit()
    best = n * args.steps / min(wall, wall_graph)
    dom = max(stats, key=lambda k: stats[k][1])
    launches, total_ms = stats[dom]
    avg_s = total_ms / launches / 1e3
    bytes_per_launch = B_ALG_G2P2G * n if dom == "g2p2g" else B_ALG_STEP * n
    achieved = bytes_per_launch / avg_s / 1e9
    out = {
        "metric": METRIC, "value": round(best, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * n / best, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "impact_ensemble_256: batched perfect-plastic block impacts",
                   "instances": len(scenes), "particles": n, "seed": W.SEED,
                   "autotuned_persist_wgs": tuned["persist_wgs"] if tuned else None,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_us": round(avg_s * 1e6, 2), "alg_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": None,
        "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in stats.items()},
        "audit_nonfinite": audit["nonfinite"],
    }
    print(json.dumps(out))
    ctx.close()
    return 0


def bench_gk_disk(args) -> int:
    """Brazilian-disk compression with Grady-Kipp damage (1,266,720 particles: disk plus two
    rigid platens, two-body contact step, FP64 eig3 per stressed particle)."""
    from graphyt_amd import workloads as W
    return bench_contact(args, W.gk_disk(), "gk_disk: Brazilian disk R 0.05 t 0.02, h 0.001, rigid platens +-0.1 m/s",
                         "f32 (eig3 and damage f64)")


def bench_plane_brazilian(args) -> int:
    """The plane contact step at about a million particles: the Brazilian test of
    Example_Scripts/brazillian2D.py (Drucker-Prager rock with Grady-Kipp damage between two
    rigid platens) at h = 0.125 mm, 1601 x 1601 nodes."""
    from graphyt_amd import workloads as W
    return bench_contact(args, W.brazilian_disk_2d(h=0.000125, v=2.0, jitter=0.1),
                         "plane_brazilian: 2D Brazilian disk R 0.04, h 0.125 mm, DP + Grady-Kipp, rigid platens +-2 m/s",
                         "f32 (damage f64)")


def bench_contact(args, sc, desc, dtype) -> int:
    """A contact workload: per-kernel HIP-event timing, then the captured hipGraph."""
    from graphyt_amd import runtime
    log(f"workload {sc.name}: {sc.n} particles, nodes {sc.dims}, dt {sc.resolved_dt():.4e}")
    ctx = runtime.Context(sc, graph=True)
    ctx.run(args.warmup)
    ctx.synchronize()
    ctx.set_option(runtime.OPT_PROFILE, 1)
    ctx.reset_kernel_stats()
    t0 = time.perf_counter()
    ctx.run(args.steps, step0=args.warmup)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    stats = ctx.kernel_stats()
    ctx.set_option(runtime.OPT_PROFILE, 0)
    t0 = time.perf_counter()
    ctx.run(args.steps, step0=args.warmup + args.steps)
    ctx.synchronize()
    wall_graph = time.perf_counter() - t0
    audit = ctx.audit()
    best = sc.n * args.steps / min(wall, wall_graph)
    out = {
        "metric": METRIC, "value": round(best, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * sc.n / best, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": desc, "particles": sc.n, "nodes": list(sc.dims), "dt": sc.resolved_dt(),
                   "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
        "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in stats.items()},
        "audit_nonfinite": audit["nonfinite"],
    }
    print(json.dumps(out))
    ctx.close()
    return 0


# Plane dam break (Newtonian, single body): the fused plane kernel reads and writes 22
# state words per particle (q 3, v 3, sigma 6, eps 6, rho, m, tag, pid: 176 B), reads its
# list entry and writes key and rank (12 B), and per tile of 8 x 8 cells (256 particles when
# full) stages 100 node velocities and writes 100 node partials (32 B each): 25 B per
# particle. The step adds place (12 B) and the grid update (about 1/4 node per particle,
# partial reads and a VEL write: 20 B).
B_ALG_PLANE_G2P2G = 176 + 12 + 25
B_ALG_PLANE_STEP = B_ALG_PLANE_G2P2G + 12 + 20


def plane_scene():
    from graphyt_amd import workloads as W
    return W.dam_break_2d(h=0.00125)


def bench_plane(args) -> int:
    """The plane (is3D = false) path at about a million particles: the dam break of
    Example_Scripts/dambreak_2D.py at h = 1.25 mm (801 x 801 nodes, 944,860 particles)."""
    from graphyt_amd import runtime
    sc = plane_scene()
    log(f"workload plane_dam_break: {sc.n} particles, nodes {sc.dims}, dt {sc.resolved_dt():.4e}")
    ctx = runtime.Context(sc, graph=True)
    ctx.run(args.warmup)
    ctx.synchronize()
    ctx.set_option(runtime.OPT_PROFILE, 1)
    ctx.reset_kernel_stats()
    t0 = time.perf_counter()
    ctx.run(args.steps, step0=args.warmup)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    stats = ctx.kernel_stats()
    ctx.set_option(runtime.OPT_PROFILE, 0)
    t0 = time.perf_counter()
    ctx.run(args.steps, step0=args.warmup + args.steps)
    ctx.synchronize()
    wall_graph = time.perf_counter() - t0
    audit = ctx.audit()
    log("kernel stats (launches, total ms):", stats)
    best = sc.n * args.steps / min(wall, wall_graph)
    launches, total_ms = stats["g2p2g"]
    avg_s = total_ms / launches / 1e3
    bpl = B_ALG_PLANE_G2P2G * sc.n
    achieved = bpl / avg_s / 1e9
    cpu = None if args.no_cpu else run_cpu_baseline(min(args.cpu_steps, 30), "plane_dam_break")
    out = {
        "metric": "particle-steps/s (plane GIMP, 0.94M pts)", "value": round(best, 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * sc.n / best, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plane_dam_break: 2D Newtonian dam break (dambreak_2D.py), h 1.25 mm",
                   "particles": sc.n, "nodes": list(sc.dims), "dt": sc.resolved_dt(), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "g2p2g", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("g2p2g", PLANE_TRAFFIC_PROFILE),
                     "avg_launch_us": round(avg_s * 1e6, 2), "alg_bytes_per_launch": bpl},
        "cpu_baseline": cpu,
        "step_hbm_frac": round(B_ALG_PLANE_STEP * best / (HBM_PEAK_GBS * 1e9), 4),
        "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in stats.items()},
        "audit_nonfinite": audit["nonfinite"],
    }
    print(json.dumps(out))
    ctx.close()
    return 0


def surrogate_config(args) -> dict:
    """The learned Drucker-Prager return map at scale. Training: constitutive samples of
    the plain GPU step on the 1,024,000-particle column at friction angles 25, 30, 35 and
    40 degrees (5 pairs of consecutive states each: 20.5M samples), the network trained on
    the GPU. Map level: a held-out pair of the column at 32.5 degrees (step 600) through
    the closed form, the reference's 201-pass loop and the gated network. Step level: the
    4,194,304-particle column (h = 0.00625) stepped with each of the three multipliers;
    stress of the learned and loop steps against the closed-form step after the same
    steps."""
    import gc
    import numpy as np
    from graphyt_amd import runtime, surrogate as S, workloads as W

    def rel(a, b):
        return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))

    t0 = time.perf_counter()
    data = S.collect_columns(lambda phi: W.dp_column(phi=phi), (25.0, 30.0, 35.0, 40.0), pairs=5, gap=150, start=1)
    n_samples = int(sum(len(x) for _, x in data))
    t_collect = time.perf_counter() - t0
    sur = S.Surrogate(seed=0)
    t1 = time.perf_counter()
    st = sur.fit(data, epochs=args.sur_epochs)
    t_fit = time.perf_counter() - t1
    del data
    gc.collect()
    log(f"surrogate: {n_samples} samples in {t_collect:.1f} s, trained in {t_fit:.1f} s: {st}")
    held_sc = W.dp_column(phi=32.5)
    held = S.collect(held_sc, pairs=1, gap=1, start=600)
    ev = S.evaluate(sur, held_sc.mats[0], held, reps=5)
    del held
    log("surrogate map level:", ev)
    steps, warm = args.sur_steps, 20
    runs, final = {}, {}
    for name, mode in (("closed_form", runtime.DP_CLOSED), ("reference_loop", runtime.DP_REFERENCE),
                       ("learned", runtime.DP_LEARNED)):
        sc = W.dp_column_4m(dp_return=mode)
        ctx = runtime.Context(sc, graph=True)
        if mode == runtime.DP_LEARNED:
            ctx.set_return_network(sur.net)
        ctx.run(warm, step0=0)
        ctx.synchronize()
        ta = time.perf_counter()
        ctx.run(steps, step0=warm)
        ctx.synchronize()
        ms = (time.perf_counter() - ta) * 1e3 / steps
        final[name] = ctx.download(["pos", "vel", "stress"])
        runs[name] = {"ms_per_step": round(ms, 4), "particle_steps_per_s": round(sc.n / ms * 1e3, 1)}
        if mode == runtime.DP_LEARNED:
            runs[name]["return_map_samples"] = ctx.return_map_counts()
        n4 = sc.n
        ctx.close()
        log(f"surrogate step {name}: {runs[name]}")
    sur.close()
    cl = final["closed_form"]
    err = {name: {k: rel(final[name][k], cl[k]) for k in ("pos", "vel", "stress")}
           for name in ("learned", "reference_loop")}
    return {
        "workload": f"dp_column_4M: the Drucker-Prager column at h = 0.00625, {n4} particles, phi 30",
        "steps": steps, "warmup": warm, "step": runs,
        "speedup_vs_reference_loop": round(runs["reference_loop"]["ms_per_step"] / runs["learned"]["ms_per_step"], 4),
        "speedup_vs_closed_form": round(runs["closed_form"]["ms_per_step"] / runs["learned"]["ms_per_step"], 4),
        "rel_l2_vs_closed_form_after_steps": err,
        "map": {k: (float(f"{v:.4g}") if isinstance(v, float) else v) for k, v in ev.items()},
        "training": {"samples": n_samples, "phi_deg": [25, 30, 35, 40], "epochs": args.sur_epochs,
                     "collect_s": round(t_collect, 1), "train_s": round(t_fit, 1),
                     **{k: v for k, v in st.items() if k != "samples"}},
        "network": "14-64-64-64-1 SiLU, BF16 matrix cores (v_mfma_f32_16x16x16bf16_1k), FP32 accumulation",
    }


def bench_surrogate(args) -> int:
    """One JSON line for the learned return map at scale (surrogate_config)."""
    res = surrogate_config(args)
    out = {
        "metric": METRIC, "value": res["speedup_vs_reference_loop"],
        "unit": "x (learned-map step time / reference-loop step time, 4.2M-particle column)",
        "n_gpus": 1, "steps": res["steps"], "warmup": res["warmup"],
        "ms_per_step": res["step"]["learned"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 step, bf16 matrix-core network with f32 accumulation",
        "data": "synthetic (seeded column; training samples logged from GPU runs)",
        "config": {"workload": res["workload"], "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None, "surrogate": res,
    }
    print(json.dumps(out))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-baseline-only", action="store_true")
    ap.add_argument("--profile-json", default="")
    ap.add_argument("--workload", choices=["dp_column", "ensemble", "surrogate", "gk_disk", "plane_dam_break",
                                           "plane_brazilian"],
                    default="dp_column",
                    help="dp_column: the 1M-particle column (default); ensemble: 256 batched impacts; "
                         "surrogate: the learned return map at 4.2M particles alone")
    ap.add_argument("--autotune", action="store_true", help="run gy_autotune before timing")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="leave out the learned return map at 4.2M particles (the `surrogate` object)")
    ap.add_argument("--sur-steps", type=int, default=1000, help="timed steps per return-map mode at 4.2M")
    ap.add_argument("--sur-epochs", type=int, default=8, help="training epochs of the return-map network")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()

    if args.calibrate_copy:
        return copy_calibration()

    if args.cpu_baseline_only:
        print(json.dumps(cpu_baseline(args.cpu_steps, 2, args.workload)))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the MPM step runs on one GPU; multi-GPU runs (slabs along x with a "
              "3-plane overlap exchange, or instances spread over GPUs) are not built", file=sys.stderr)
        return 2

    from graphyt_amd import runtime, workloads as W
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "surrogate":
        return bench_surrogate(args)
    if args.workload == "gk_disk":
        return bench_gk_disk(args)
    if args.workload == "plane_brazilian":
        return bench_plane_brazilian(args)
    if args.workload == "plane_dam_break":
        return bench_plane(args)
    sc = W.dp_column()
    log(f"workload dp_column: {sc.n} particles, nodes {sc.dims}, dt {sc.resolved_dt():.4e}")
    ctx = runtime.Context(sc, graph=True)
    if args.autotune:
        log("autotune:", ctx.autotune(10))
    ctx.run(args.warmup, step0=0)
    ctx.synchronize()

    # timed region: per-kernel HIP events on the context stream
    ctx.set_option(runtime.OPT_PROFILE, 1)
    ctx.reset_kernel_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(args.steps, step0=args.warmup)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    stats = ctx.kernel_stats()
    ctx.set_option(runtime.OPT_PROFILE, 0)

    # the same steps replayed through the captured hipGraph (production path)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(args.steps, step0=args.warmup + args.steps)
    ctx.synchronize()
    wall_graph = time.perf_counter() - t0
    audit = ctx.audit()
    log("audit after", args.warmup + 2 * args.steps, "steps:", audit)
    log("kernel stats (launches, total ms):", stats)

    value = sc.n * args.steps / wall
    value_graph = sc.n * args.steps / wall_graph
    dom = max(stats, key=lambda k: stats[k][1])
    launches, total_ms = stats[dom]
    avg_s = total_ms / launches / 1e3
    bytes_per_launch = (B_ALG_G2P2G if dom == "g2p2g" else B_ALG_STEP) * sc.n
    achieved = bytes_per_launch / avg_s / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
                "avg_launch_us": round(avg_s * 1e6, 2), "alg_bytes_per_launch": bytes_per_launch}

    cpu = None if args.no_cpu else run_cpu_baseline(args.cpu_steps)
    ctx.close()
    surrogate = None
    if not args.no_surrogate:
        try:
            surrogate = surrogate_config(args)
        except Exception as e:  # reported in the line, the column's figures stand
            surrogate = {"error": f"{type(e).__name__}: {e}"}

    best = max(value, value_graph)
    out = {
        "metric": METRIC, "value": round(best, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * sc.n / best, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "dp_column_1M: 3D Drucker-Prager granular column collapse",
                   "particles": sc.n, "nodes": list(sc.dims), "cellsize": sc.h, "psep": sc.psep,
                   "material": "model 4, rho 2000, K 5e7, G 2e7, c 0, phi 30, psi 0 (guarded apex)",
                   "dt": sc.resolved_dt(), "seed": W.SEED, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "value_events": round(value, 1), "value_graph": round(value_graph, 1),
        "step_hbm_frac": round(B_ALG_STEP * best / (HBM_PEAK_GBS * 1e9), 4),
        "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in stats.items()},
        "audit_nonfinite": audit["nonfinite"],
        "surrogate": surrogate,
    }
    if args.profile_json:
        with open(args.profile_json, "w") as f:
            json.dump({"stats": stats, "out": out}, f, indent=1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
