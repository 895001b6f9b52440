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
