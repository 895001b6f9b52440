"""The batched ensemble (gy_create_batch / gy_upload_batch / gy_download_inst) and the
autotuner, through the C ABI: every instance against the FP64 oracle run on its own,
batched against one-instance contexts, and the autotuner's state restore."""
import numpy as np
import pytest

from graphyt_amd import runtime, workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ["pos", "vel", "stress", "acc_plastic_strain"]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def small_ensemble(count=6, seed=7):
    return W.impact_ensemble(count=count, nmin=500, nmax=4000, seed=seed)


def test_each_instance_matches_its_own_oracle_run(oracle_mod):
    scenes = small_ensemble()
    b = runtime.BatchContext(scenes)
    b.run(150)
    for i, sc in enumerate(scenes):
        o = oracle_mod.CpuSim(sc, "restatement")
        assert o.dt == pytest.approx(b.dt_of(i), rel=1e-15)
        o.run(150)
        ref, got = o.get(FIELDS), b.download(i, FIELDS)
        r = {k: rel(got[k], ref[k]) for k in ("pos", "vel", "stress")}
        print(f"instance {i} ({sc.n} particles): " + ", ".join(f"{k} {v:.2e}" for k, v in r.items()), flush=True)
        assert r["pos"] <= 1e-6, i
        assert r["vel"] <= 1e-4, i
        assert r["stress"] <= 1e-4, i
    b.close()


def test_batched_equals_separate_contexts():
    scenes = small_ensemble(count=4, seed=11)
    b = runtime.BatchContext(scenes)
    b.run(80)
    for i, sc in enumerate(scenes):
        g = runtime.Context(sc)
        g.run(80)
        one, many = g.download(FIELDS), b.download(i, FIELDS)
        assert rel(many["pos"], one["pos"]) <= 1e-7
        assert rel(many["stress"], one["stress"]) <= 1e-4
        g.close()
    b.close()


def test_full_ensemble_batched_equals_separate_contexts():
    """The batched-ensemble configuration at its full size (256 impacts, 4.9M particles) in
    one batched context: every instance finite with its mass unchanged, and eight instances
    across the size range against one-instance contexts of the same scenes (the instances
    share no state; batching changes only the order of FP64 LDS atomics)."""
    scenes = W.impact_ensemble()
    steps = 100
    b = runtime.BatchContext(scenes)
    b.run(steps)
    sizes = np.array([sc.n for sc in scenes])
    print(f"{len(scenes)} instances, {int(sizes.sum())} particles", flush=True)
    for i, sc in enumerate(scenes):
        got = b.download(i, ["pos", "stress", "mass"])
        assert all(np.isfinite(got[k]).all() for k in got), i
        # m = rho * psep^3 per particle, set at the upload (Solver.cpp:38-60), never changed
        assert np.allclose(got["mass"], sc.mats[0].density * sc.psep ** 3, rtol=1e-6, atol=0), i
    order = np.argsort(sizes)
    picks = sorted(set(int(order[k]) for k in np.linspace(0, len(scenes) - 1, 8).round().astype(int)))
    for i in picks:
        g = runtime.Context(scenes[i])
        g.run(steps)
        one, many = g.download(FIELDS), b.download(i, FIELDS)
        g.close()
        r = {k: rel(many[k], one[k]) for k in ("pos", "vel", "stress")}
        print(f"instance {i} ({scenes[i].n} particles): batched vs alone " +
              ", ".join(f"{k} {v:.2e}" for k, v in r.items()), flush=True)
        assert r["pos"] <= 1e-7 and r["vel"] <= 1e-4 and r["stress"] <= 1e-4, i
    b.close()


def test_autotune_restores_the_particle_state():
    scenes = small_ensemble(count=5, seed=3)
    b = runtime.BatchContext(scenes)
    b.run(20)
    before = [b.download(i, ["pos", "vel", "stress"]) for i in range(len(scenes))]
    res = b.autotune(trial_steps=5)
    assert res["persist_wgs"] % 8 == 0 and res["ms_per_step"] > 0
    after = [b.download(i, ["pos", "vel", "stress"]) for i in range(len(scenes))]
    for x, y in zip(before, after):
        for k in x:
            assert np.array_equal(x[k], y[k])
    b.run(20, step0=20)
    ref = runtime.BatchContext(scenes)
    ref.run(40)
    for i in range(len(scenes)):
        assert rel(b.download(i, ["pos"])["pos"], ref.download(i, ["pos"])["pos"]) <= 1e-7
    b.close()
    ref.close()


def test_autotune_shapes_cache_is_reused(tmp_path):
    """gy_autotune_shapes tunes the workgroup count, the fused kernel's build and the grid
    kernel's workgroups; the winner goes to the cache file under the workload's shape, and
    a second context of the same shape takes it from there without trials. Tuning leaves
    the state as it was (bit for bit) and the tuned launches give the default's results."""
    cache = str(tmp_path / "tune.txt")
    sc = W.dp_column_small(steps=60)
    a = runtime.Context(sc)
    a.run(10)
    before = a.download(["pos", "vel", "stress"])
    t1 = a.autotune_shapes(trial_steps=5, cache_path=cache)
    print("tuned:", t1, flush=True)
    assert not t1["from_cache"] and t1["persist_wgs"] % 8 == 0 and t1["ms_per_step"] > 0
    assert t1["fused_waves"] in (2, 3) and t1["grid_wgs_factor"] in (1, 2, 4)
    assert a.tuning()["persist_wgs"] == t1["persist_wgs"]
    after = a.download(["pos", "vel", "stress"])
    for k in before:
        assert np.array_equal(before[k], after[k])
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and lines[0].startswith("v1|")
    a.run(40, step0=10)
    tuned = a.download(["pos", "vel", "stress"])
    a.close()
    b = runtime.Context(sc)
    t2 = b.autotune_shapes(trial_steps=5, cache_path=cache)
    assert t2["from_cache"]
    for k in ("persist_wgs", "fused_waves", "grid_wgs_factor"):
        assert t2[k] == t1[k]
    b.close()
    ref = runtime.Context(sc)
    ref.run(50)
    r = ref.download(["pos", "vel", "stress"])
    ref.close()
    assert rel(tuned["pos"], r["pos"]) <= 1e-9
    assert rel(tuned["vel"], r["vel"]) <= 1e-6 and rel(tuned["stress"], r["stress"]) <= 1e-5
    # another shape (more particles) is not served from this entry
    c = runtime.Context(W.dp_column_mid(steps=10))
    assert not c.autotune_shapes(trial_steps=2, cache_path=cache)["from_cache"]
    c.close()
    assert len(open(cache).read().splitlines()) == 2
