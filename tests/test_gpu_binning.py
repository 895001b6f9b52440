"""Counting sort by cell on the GPU, bit-exact against a CPU recount (numpy bincount,
cumsum and stable argsort). Calls go through the C ABI (gy_counting_sort,
gy_debug_binning)."""
import numpy as np
import pytest

from graphyt_amd import runtime, workloads as W

pytestmark = pytest.mark.gpu


def recount(keys, nkeys):
    counts = np.bincount(keys, minlength=nkeys).astype(np.int32)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int32)
    order = np.argsort(keys, kind="stable").astype(np.int32)
    return counts, offsets, order


@pytest.mark.parametrize("n,nkeys,seed", [(1, 1, 0), (7, 3, 1), (1000, 64, 2), (100_000, 2_000_000, 3),
                                          (300_000, 50_000, 4), (5000, 4096 * 3 + 17, 5)])
def test_counting_sort_bit_exact(n, nkeys, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, nkeys, size=n).astype(np.int32)
    got = runtime.counting_sort(keys, nkeys)
    for g, e in zip(got, recount(keys, nkeys)):
        assert np.array_equal(g, e)


def test_counting_sort_collisions_and_sorted_input():
    keys = np.zeros(70_000, np.int32)          # every key collides
    keys[::3] = 5
    for k in (keys, np.sort(keys), np.sort(keys)[::-1].copy()):
        got = runtime.counting_sort(k, 6)
        for g, e in zip(got, recount(k, 6)):
            assert np.array_equal(g, e)


def test_counting_sort_rejects_bad_keys():
    with pytest.raises(runtime.GyError):
        runtime.counting_sort(np.array([0, 3], np.int32), 3)


def test_step_binning_matches_recount_after_steps():
    sc = W.dp_column_small()
    ctx = runtime.Context(sc, graph=False)
    for nsteps in (0, 7):
        if nsteps:
            ctx.run(nsteps)
        b = ctx.binning()
        counts, offsets, order = recount(b["keys"], len(b["counts"]))
        assert np.array_equal(b["counts"], counts)
        assert np.array_equal(b["offsets"], offsets)
        assert np.array_equal(b["order"], order)
        assert counts.sum() == sc.n
    ctx.close()
