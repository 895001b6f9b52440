"""The CPU checker of the return-map network pinned by mathematics: bfloat16 rounding on
known values, and its back-propagation against central finite differences of its own
loss (float64, no rounding)."""
import numpy as np

import mlp_ref as R


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.0, 0.1], np.float32)
    got = R.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1.0); 1 + 1.5*2^-8 rounds up
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0078125 and got[3] == 1.0078125
    assert got[4] == -3.0 and abs(got[5] - 0.1) < 0.1 * 2 ** -8


def test_layout_sizes():
    assert R.NPARAM == 10448 and R.NW == 16 * 64 + 2 * 64 * 64 + 64 * 16


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = rng.normal(0, 0.3, R.NPARAM)
    x = rng.normal(0, 1, (9, R.NIN))
    t = rng.normal(0, 1, (9, R.NOUT))
    g, _ = R.gradient(p, x, t, bf16_ops=False)
    h = 1e-6
    # padded weights that see no input/output have zero gradient
    Ws, _ = R.unpack(g)
    assert np.all(Ws[0][R.NIN:, :] == 0) and np.all(Ws[3][:, R.NOUT:] == 0)
    idx = np.r_[rng.choice(R.NW, 40, replace=False), R.NW + rng.choice(R.NPARAM - R.NW - 9, 20, replace=False)]
    for i in idx:
        e = np.zeros(R.NPARAM)
        e[i] = h
        lp = R.gradient(p + e, x, t, bf16_ops=False)[1]
        lm = R.gradient(p - e, x, t, bf16_ops=False)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
