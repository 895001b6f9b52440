"""The learned return map on the GPU: BF16 matrix-core forward and backward against the
CPU checker (tests/mlp_ref.py, which rounds to bfloat16 where the kernels do), training,
the gate (elastic and out-of-range samples take the exact map, bit for bit), and the
trained surrogate on Drucker-Prager column data against the exact device map and the
FP64 oracle."""
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
    x = rng.normal(0, 1, (300, 13)).astype(np.float32)   # 300: a partial last 64-row tile
    assert rel(net.forward(x), R.forward(p, x)) <= 1e-5
    net.close()


def test_gradient_matches_cpu_checker():
    rng = np.random.default_rng(2)
    net = runtime.Mlp(4)
    p = net.params.astype(np.float64)
    x = rng.normal(0, 1, (200, 13)).astype(np.float32)
    t = rng.normal(0, 0.5, (200, 7)).astype(np.float32)
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
    x = rng.normal(0, 1, (40000, 13)).astype(np.float32)
    t = R.forward(R.pack(Ws, bs), x, bf16_ops=False).astype(np.float32)
    net = runtime.Mlp(9)
    _, loss0 = net.gradient(x[:4096], t[:4096])
    net.train(x, t, epochs=15, batch=1024, lr=2e-3, lr_decay=0.9)
    _, loss1 = net.gradient(x[:4096], t[:4096])
    assert loss1 < 0.1 * loss0
    net.close()


def test_gate_sends_elastic_and_out_of_range_samples_to_the_exact_map():
    mat = S.column_material()
    net = runtime.Mlp(5)
    mean, std = np.zeros(13, np.float32), np.ones(13, np.float32)
    net.set_norm(mean, std, np.full(13, -1.0, np.float32), np.full(13, 1.0, np.float32), np.ones(7, np.float32))
    rng = np.random.default_rng(6)
    n = 5000
    x = np.zeros((n, 13), np.float32)
    x[:, 6:9] = -rng.uniform(1e3, 1e4, (n, 1))          # hydrostatic compression: elastic
    x[:, 0:3] = -rng.uniform(0, 1e-7, (n, 3))
    ex = runtime.dp_map(mat, x)
    su = runtime.dp_map(mat, x, net)
    _, yld = S.trial(mat, x)
    assert not yld.any() and su["network"] == 0 and su["fallback"] == 0
    assert np.array_equal(su["out"], ex["out"])
    # strongly sheared samples yield; with a [-1, 1] trained range they are out of range
    y = x.copy()
    y[:, 9] = 5e4
    _, yld = S.trial(mat, y)
    assert yld.all()
    ex, su = runtime.dp_map(mat, y), runtime.dp_map(mat, y, net)
    assert su["network"] == 0 and su["fallback"] == n
    assert np.array_equal(su["out"], ex["out"])
    net.close()


def test_surrogate_on_column_data(oracle_mod):
    sc = W.dp_column_mid()
    mat = sc.mats[0]
    train = S.collect(sc, pairs=12, gap=20, start=20)
    held = S.collect(sc, pairs=2, gap=30, start=400)
    sur = S.Surrogate(mat, seed=0)
    st = sur.fit(train, epochs=20)
    assert st["yielding"] > 1000
    ev = S.evaluate(sur, held)
    print("surrogate on held-out column samples:", ev)
    # the gate sends every yielding sample somewhere and leaves elastic ones exact
    assert abs(ev["network"] + ev["fallback"] - ev["yielding"]) <= 0.001 * ev["samples"]
    assert ev["stress_rel_l2_yielding"] <= 0.25 and ev["epbar_rel_l2_yielding"] <= 0.05
    # against the FP64 oracle's return map on a subset of yielding samples: the exact
    # FP32 device map agrees to FP32 rounding, the surrogate to its fitted accuracy
    x = np.asarray(held, np.float32)
    _, yld = S.trial(mat, x)
    idx = np.flatnonzero(yld)[:300]
    ref = np.array([oracle_mod.constitutive(mat, x[i, :6].astype(np.float64), x[i, 6:12].astype(np.float64), 1.0,
                                            epbar=float(x[i, 12]))[0] for i in idx])
    exact = runtime.dp_map(mat, x[idx])["out"][:, :6]
    surr = sur.apply(x[idx])["out"][:, :6]
    assert rel(exact, ref) <= 1e-4
    assert rel(surr, ref) <= 0.25
    sur.close()
