"""gy_counting_sort_noalloc, the zero-allocation counting sort, through the C ABI: device
inputs and outputs allocated by the caller (here with hipMalloc from the HIP runtime the
library links), the library's own device scratch, two calls in a row (the second
smaller), each against gy_counting_sort (the allocating entry point) and the numpy closed
form: bit-exact."""
import ctypes as C

import numpy as np
import pytest

from graphyt_amd import runtime

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    runtime.load()  # the library has the runtime loaded: this is the same instance
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _dev(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), nbytes) == 0
    return p


def _noalloc(lib, hip, keys_np, nkeys):
    n = keys_np.size
    bufs = [_dev(hip, 4 * n), _dev(hip, 4 * nkeys), _dev(hip, 4 * nkeys), _dev(hip, 4 * n)]
    keys, counts, offsets, order = bufs
    try:
        assert hip.hipMemcpy(keys, keys_np.ctypes.data, 4 * n, H2D) == 0
        fill = np.full(max(n, nkeys), -7, np.int32)
        for p, m in ((counts, nkeys), (offsets, nkeys), (order, n)):
            assert hip.hipMemcpy(p, fill.ctypes.data, 4 * m, H2D) == 0
        assert lib.gy_counting_sort_noalloc(keys, n, nkeys, counts, offsets, order, None) == 0
        assert hip.hipDeviceSynchronize() == 0
        out = [np.empty(nkeys, np.int32), np.empty(nkeys, np.int32), np.empty(n, np.int32)]
        for host, p in zip(out, (counts, offsets, order)):
            assert hip.hipMemcpy(host.ctypes.data, p, host.nbytes, D2H) == 0
        return out
    finally:
        for p in bufs:
            hip.hipFree(p)


def test_noalloc_sort_matches_the_allocating_sort_twice():
    lib, hip = runtime.load(), _hip()
    rng = np.random.default_rng(20261203)
    for n, nkeys in ((1_000_000, 700_000), (123_457, 4_099)):  # the second call smaller
        print(f"no-alloc counting sort: n {n}, nkeys {nkeys}", flush=True)
        keys = rng.integers(0, nkeys, size=n, dtype=np.int32)
        keys[: n // 3] = rng.integers(0, max(1, nkeys // 50), size=n // 3, dtype=np.int32)  # crowded keys
        c0, o0, r0 = runtime.counting_sort(keys, nkeys)
        c1, o1, r1 = _noalloc(lib, hip, keys, nkeys)
        assert np.array_equal(c1, c0) and np.array_equal(o1, o0) and np.array_equal(r1, r0)
        assert np.array_equal(c1, np.bincount(keys, minlength=nkeys))
        assert np.array_equal(r1, np.argsort(keys, kind="stable"))


def test_noalloc_sort_rejects_sizes_beyond_its_scratch():
    lib = runtime.load()
    assert lib.gy_counting_sort_noalloc(C.c_void_p(1), (1 << 20) + 1, 16, None, None, None, None) == 1
    assert lib.gy_counting_sort_noalloc(C.c_void_p(1), 16, (1 << 20) + 1, None, None, None, None) == 1
