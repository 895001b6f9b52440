"""The learned return map on the GPU: BF16 matrix-core forward and backward against the
CPU checker (tests/mlp_ref.py, which rounds to bfloat16 where the kernels do), training,
the step's wavefront evaluation against the training forward, the gate (elastic and
out-of-range samples take the exact map, bit for bit), the reference's loop against the
closed form, and the trained surrogate on Drucker-Prager column data against the exact
device map and the FP64 oracle."""
import numpy as np
import pytest

import mlp_ref as R
from graphyt_amd import runtime, surrogate as S, workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_forward_matches_cpu_checker():
    rng = np.random.default_rng(1)
    net = runtime.Mlp(3)
    p = net.params.astype(np.float64)
    x = rng.normal(0, 1, (300, R.NIN)).astype(np.float32)   # 300: a partial last 64-row tile
    got, want = net.forward(x), R.forward(p, x)
    # most outputs agree to FP32 rounding; an activation whose FP32 sum lands on the other
    # side of a BF16 rounding boundary (a different accumulation order) moves its sample's
    # one output by ~1e-4 relative
    assert np.median(np.abs(got - want)) <= 1e-6 * np.abs(want).max()
    assert rel(got, want) <= 1e-4
    net.close()


def test_gradient_matches_cpu_checker():
    rng = np.random.default_rng(2)
    net = runtime.Mlp(4)
    p = net.params.astype(np.float64)
    x = rng.normal(0, 1, (200, R.NIN)).astype(np.float32)
    t = rng.normal(0, 0.5, (200, R.NOUT)).astype(np.float32)
    g, loss = net.gradient(x, t)
    gr, lr_ = R.gradient(p, x, t)
    assert loss == pytest.approx(lr_, rel=1e-5)
    assert rel(g, gr) <= 1e-4
    assert float(g @ gr) / np.linalg.norm(g) / np.linalg.norm(gr) >= 0.99999
    # every layer's block of the gradient is checked, not only the largest
    o = 0
    for a, b in R.SHAPES:
        blk = slice(o, o + a * b)
        if np.linalg.norm(gr[blk]) > 0:
            assert rel(g[blk], gr[blk]) <= 1e-3
        o += a * b
    net.close()


def test_training_fits_a_teacher_network():
    rng = np.random.default_rng(3)
    teacher = R.unpack(rng.normal(0, 0.25, R.NPARAM))
    Ws, bs = teacher
    Ws[0][R.NIN:, :] = 0
    Ws[3][:, R.NOUT:] = 0
    x = rng.normal(0, 1, (40000, R.NIN)).astype(np.float32)
    t = R.forward(R.pack(Ws, bs), x, bf16_ops=False).astype(np.float32)
    net = runtime.Mlp(9)
    _, loss0 = net.gradient(x[:4096], t[:4096])
    net.train(x, t, epochs=15, batch=1024, lr=2e-3, lr_decay=0.9)
    _, loss1 = net.gradient(x[:4096], t[:4096])
    assert loss1 < 0.1 * loss0
    net.close()


def yielding_samples(mat, n, seed):
    """Raw samples whose trial states yield: compression with strong shear."""
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 13), np.float32)
    x[:, 0:6] = rng.normal(0, 2e-5, (n, 6))
    x[:, 6:9] = -1e4 + rng.normal(0, 3e3, (n, 3))
    x[:, 9:12] = rng.normal(0, 6e3, (n, 3))
    x[:, 12] = rng.uniform(0, 0.05, n)
    _, yld, _ = S.trial(mat, x)
    return x[yld]


def test_step_evaluation_matches_the_training_forward():
    """The wavefront evaluation of the map kernel and the step (register-resident
    activations, fragment-ordered weights, hardware exp2/reciprocal SiLU) against the
    training kernels' forward on the same normalised inputs: a transposed fragment, a
    wrong lane of the cross-lane transposes or a wrong bias offset gives O(1) errors.
    Output scaled as 1 + 0.1 y, accepted in [0.5, 1.5]; the samples' overshoot is small
    enough that the return's branch is the cone one across that range, and large enough
    that the device's and numpy's roundings of F agree to 1e-6."""
    mat = S.column_material()
    x = yielding_samples(mat, 12000, 11)
    _, _, sJ2, F, _ = S._trial(mat, x)
    x = x[(F < 0.6 * sJ2) & (F > 0.05 * sJ2)]   # F well above the rounding of sqrt(J2) + eta I1
    f = S.features(mat, x)
    mean, std = f.mean(axis=0), np.maximum(f.std(axis=0), 1.0).astype(np.float32)
    net = runtime.Mlp(21)
    big = np.full(R.NIN, 1e30, np.float32)
    net.set_norm(mean, std, -big, big, 0.1, 1.0, 0.5, 1.5)
    xn = ((f - mean) / std).astype(np.float32)
    y = net.forward(xn)[:, 0].astype(np.float64)
    r = runtime.dp_map(mat, x, net)
    _, _, F = S.trial(mat, x)
    got = (r["multiplier"] / S.linear_multiplier(mat, F) - 1.0) / 0.1
    inside = np.abs(y) < 4.9
    print("samples", len(x), "network", r["network"], "outputs inside the range", int(inside.sum()),
          "rel L2", rel(got[inside], y[inside]), flush=True)
    assert inside.mean() > 0.99 and abs(r["network"] - inside.sum()) <= 0.001 * len(x)
    assert rel(got[inside], y[inside]) <= 2e-3
    net.close()


def test_gate_sends_elastic_and_out_of_range_samples_to_the_exact_map():
    mat = S.column_material()
    net = runtime.Mlp(5)
    mean, std = np.zeros(R.NIN, np.float32), np.ones(R.NIN, np.float32)
    std[13] = 100.0   # friction angle 30 -> 0.3, inside [-1, 1]
    net.set_norm(mean, std, np.full(R.NIN, -1.0, np.float32), np.full(R.NIN, 1.0, np.float32), 1.0, 1.0, 0.5, 1.5)
    rng = np.random.default_rng(6)
    n = 5000
    x = np.zeros((n, 13), np.float32)
    x[:, 6:9] = -rng.uniform(1e3, 1e4, (n, 1))          # hydrostatic compression: elastic
    x[:, 0:3] = -rng.uniform(0, 1e-7, (n, 3))
    ex = runtime.dp_map(mat, x)
    su = runtime.dp_map(mat, x, net)
    _, yld, _ = S.trial(mat, x)
    assert not yld.any() and su["network"] == 0 and su["fallback"] == 0
    assert np.array_equal(su["out"], ex["out"])
    # strongly sheared samples yield; with a [-1, 1] trained range they are out of range
    y = x.copy()
    y[:, 9] = 5e4
    _, yld, _ = S.trial(mat, y)
    assert yld.all()
    ex, su = runtime.dp_map(mat, y), runtime.dp_map(mat, y, net)
    assert su["network"] == 0 and su["fallback"] == n
    assert np.array_equal(su["out"], ex["out"])
    net.close()


def test_reference_loop_map_equals_closed_form():
    """The reference's loop control (201 passes for c = 0) and the closed form reach the
    same multiplier to FP32 rounding."""
    mat = S.column_material()
    x = yielding_samples(mat, 20000, 12)
    a = runtime.dp_map(mat, x)
    b = runtime.dp_map(mat, x, mode=runtime.DP_REFERENCE)
    print("closed", a["ms"], "ms; reference loop", b["ms"], "ms", flush=True)
    assert rel(b["out"][:, :6], a["out"][:, :6]) <= 1e-6
    assert rel(b["multiplier"], a["multiplier"]) <= 1e-5


@pytest.mark.parametrize("mode", [runtime.DP_CLOSED, runtime.DP_REFERENCE])
def test_zero_deviator_in_tension_takes_the_apex(mode):
    """A tiny trial stress in hydrostatic tension (J2 underflows to 0; from about 1e-38 Pa
    down F / G underflows too) yields and returns to the apex, as in exact arithmetic,
    where dL > 0; the cone return would divide 0 by sqrt(J2) = 0. The apex stress
    I1 - K deps is 0 to rounding, or I1 itself where deps = I1 / K underflows."""
    mat = S.column_material()
    x = np.zeros((64, 13), np.float32)
    p = np.logspace(-38, -29, 64).astype(np.float32)
    x[:, 6:9] = p[:, None]
    r = runtime.dp_map(mat, x, mode=mode)
    assert np.all(np.isfinite(r["out"]))
    assert np.all(r["out"][:, 3:6] == 0.0)
    assert np.all(np.abs(r["out"][:, :3]) <= 1.0001 * p[:, None])


def test_surrogate_on_column_data(oracle_mod):
    """Trained on plain-step samples of the column at friction angles 25-40 degrees;
    evaluated on held-out samples of a later stretch of the run at 30 degrees and at an
    angle between the trained ones (32.5), against the closed-form device map and the
    FP64 oracle."""
    mk = lambda phi: W.dp_column_mid(phi=phi)
    train = S.collect_columns(mk, (25.0, 30.0, 35.0, 40.0), pairs=8, gap=40, start=1)
    sur = S.Surrogate(seed=0)
    st = sur.fit(train, epochs=20)
    print("training:", st, flush=True)
    assert st["yielding"] > 10000
    for phi in (30.0, 32.5):
        sc = W.dp_column_mid(phi=phi)
        mat = sc.mats[0]
        held = S.collect(sc, pairs=2, gap=30, start=400)
        ev = S.evaluate(sur, mat, held, reps=2)
        print(f"phi {phi}: held-out column samples:", ev, flush=True)
        # the gate sends every yielding sample somewhere and leaves elastic ones exact
        assert ev["network"] + ev["fallback"] == ev["yielding"]
        assert ev["network"] >= 0.9 * ev["yielding"]
        assert ev["stress_rel_l2_yielding"] <= 1e-3 and ev["epbar_rel_l2_yielding"] <= 1e-3
        # against the FP64 oracle's return map on a subset of yielding samples
        x = np.asarray(held, np.float32)
        _, yld, _ = S.trial(mat, x)
        idx = np.flatnonzero(yld)[:300]
        ref = np.array([oracle_mod.constitutive(mat, x[i, :6].astype(np.float64), x[i, 6:12].astype(np.float64),
                                                1.0, epbar=float(x[i, 12]))[0] for i in idx])
        exact = runtime.dp_map(mat, x[idx])["out"][:, :6]
        surr = sur.apply(mat, x[idx])["out"][:, :6]
        assert rel(exact, ref) <= 1e-4
        assert rel(surr, ref) <= 1e-3
    sur.close()
