"""The allocation-free sort entry point (nguo_lbvh_sort_pairs_noalloc) against the sort that
allocates (nguo_lbvh_sort_pairs) and numpy's stable argsort: bitwise the same pairs, on calls
in a row over device buffers, each smaller than the one before, and a clear error above the
scratch maximum. The device buffers are the test's own (hipMalloc through the HIP runtime the
library links); the entry point itself allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from conftest import nguo_hip

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip_runtime():
    nguo_hip.load()  # the library's runtime is loaded first; the same one is reused
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


class DevArray:
    def __init__(self, hip, host):
        self.hip, self.n, self.p = hip, host.nbytes, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.n, 4)) == 0
        assert hip.hipMemcpy(self.p, host.ctypes.data, self.n, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


@pytest.mark.gpu
def test_noalloc_sort_matches_allocating_sort(ctx):
    hip = hip_runtime()
    rng = np.random.default_rng(17)
    # the maximum (large tiles), then smaller inputs (small tiles, a one-tile sort); keys with
    # heavy duplication so that stability shows
    for n in (nguo_hip.NOALLOC_SORT_MAX, 300_001, 1_000):
        keys = rng.integers(0, 1 << 12, n, dtype=np.uint32)
        keys[::7] = rng.integers(0, 1 << 32, len(keys[::7]), dtype=np.uint64).astype(np.uint32)
        vals = rng.permutation(n).astype(np.uint32)
        print(f"noalloc sort n={n}", flush=True)
        want_k, want_v = nguo_hip.lbvh_sort_pairs(ctx, keys, vals)
        dk, dv = DevArray(hip, keys), DevArray(hip, vals)
        try:
            nguo_hip.lbvh_sort_pairs_noalloc(ctx, n, dk.p.value, dv.p.value)
            got_k, got_v = dk.get(np.uint32, n), dv.get(np.uint32, n)
        finally:
            dk.free()
            dv.free()
        assert np.array_equal(got_k, want_k) and np.array_equal(got_v, want_v)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(got_k, keys[order]) and np.array_equal(got_v, vals[order])


@pytest.mark.gpu
def test_noalloc_sort_rejects_beyond_maximum(ctx):
    with pytest.raises(nguo_hip.NguoError) as e:
        nguo_hip.lbvh_sort_pairs_noalloc(ctx, nguo_hip.NOALLOC_SORT_MAX + 1, 0, 0)
    assert e.value.code == nguo_hip.NGUO_E_CAPACITY and "maximum" in str(e.value)
