"""The Drucker-Prager return-map modes inside the MPM step (gy_material::dp_return) on
the column at the benchmark resolution (W.dp_column_mid: h = 0.01, 16,000 particles):
the reference's loop control against the closed form, and the learned map (the network
evaluated by the fused particle kernel on the matrix cores) against the plain GPU step and
the FP64 oracle after 1000 steps."""
import numpy as np
import pytest

from graphyt_amd import runtime, surrogate as S, workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ["pos", "vel", "stress", "acc_plastic_strain"]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def trained():
    """A network trained on plain-step samples of the column at 25-40 degrees."""
    data = S.collect_columns(lambda phi: W.dp_column_mid(phi=phi), (25.0, 30.0, 35.0, 40.0), pairs=8, gap=40,
                             start=1)
    sur = S.Surrogate(seed=0)
    print("training:", sur.fit(data, epochs=20), flush=True)
    yield sur
    sur.close()


def run(sc, steps, net=None):
    ctx = runtime.Context(sc)
    if net is not None:
        ctx.set_return_network(net)
    ctx.run(steps)
    out = ctx.download(FIELDS)
    counts = ctx.return_map_counts() if net is not None else None
    ctx.close()
    return out, counts


def test_reference_loop_step_equals_closed_form():
    n = 300
    a, _ = run(W.dp_column_mid(), n)
    b, _ = run(W.dp_column_mid(dp_return=runtime.DP_REFERENCE), n)
    r = {k: rel(b[k], a[k]) for k in FIELDS}
    print("reference loop vs closed form after", n, "steps:", r, flush=True)
    assert r["pos"] <= 1e-9 and r["vel"] <= 1e-5 and r["stress"] <= 1e-4


def test_column_4m_stays_finite():
    """The 4,194,304-particle column resting on the floor nodes: its body at rest makes
    trial stresses of ~1e-31 Pa whose J2 is below the FP32 normal range (the cone return
    divided by sqrt(J2) = 0 there, NaN from step 2)."""
    sc = W.dp_column_4m()
    ctx = runtime.Context(sc)
    ctx.run(20)
    f = ctx.download(["pos", "vel", "stress"])
    ctx.close()
    for k, v in f.items():
        assert np.all(np.isfinite(v)), k


def test_learned_step_needs_a_network():
    ctx = runtime.Context(W.dp_column_small(dp_return=runtime.DP_LEARNED))
    with pytest.raises(runtime.GyError) as e:
        ctx.run(1)
    assert e.value.code == 5   # GY_ERR_STATE
    ctx.close()


def test_modes_other_than_closed_form_need_a_3d_single_body_context():
    m = S.column_material(dp_return=runtime.DP_REFERENCE)
    with pytest.raises(runtime.GyError) as e:
        runtime.Context(W.square_2d(mat=m))
    assert e.value.code == 2   # GY_ERR_UNSUPPORTED


@pytest.mark.parametrize("phi", [30.0, 35.0])
def test_learned_step_against_plain_step_and_oracle(trained, oracle_mod, phi):
    steps = 1000
    got, counts = run(W.dp_column_mid(phi=phi, dp_return=runtime.DP_LEARNED), steps, trained.net)
    plain, _ = run(W.dp_column_mid(phi=phi), steps)
    o = oracle_mod.CpuSim(W.dp_column_mid(phi=phi), "restatement")
    o.run(steps)
    ref = o.get(FIELDS)
    vs_plain = {k: rel(got[k], plain[k]) for k in FIELDS}
    vs_oracle = {k: rel(got[k], ref[k]) for k in FIELDS}
    plain_vs_oracle = {k: rel(plain[k], ref[k]) for k in FIELDS}
    print(f"phi {phi}: return-map samples {counts}", flush=True)
    print("learned vs plain GPU:", vs_plain, flush=True)
    print("learned vs FP64 oracle:", vs_oracle, flush=True)
    print("plain GPU vs FP64 oracle:", plain_vs_oracle, flush=True)
    # yielding samples the gate does not accept take the closed form
    assert counts["network"] > 100 * steps and counts["fallback"] <= 0.05 * counts["network"]
    for k in ("pos", "vel", "stress"):
        assert np.all(np.isfinite(got[k]))
    # measured: stress 3.5e-5 (plain) and 4.4e-5 (oracle) at 30 and 35 degrees, as close as
    # the plain GPU step itself is to the oracle (2.9e-5); held to the step's 1e-4 bar
    assert vs_plain["stress"] <= 1e-4 and vs_oracle["stress"] <= 1e-4
    assert vs_plain["vel"] <= 1e-4 and vs_oracle["vel"] <= 1e-4
    assert vs_plain["pos"] <= 1e-6 and vs_oracle["pos"] <= 1e-6
