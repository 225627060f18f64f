"""hipcstone_sort_by_key_u64_noalloc (static device scratch, no allocation) against hipcstone_sort_by_key_u64 (caller
workspace), called through the C ABI: two calls in a row on the same device, the second with a smaller input, each
compared bit for bit (keys and carried values) with the workspace entry on the same input. The second call reuses the
scratch the first one left behind, so anything that must start at zero has to be cleared by the call itself.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sort_noalloc(hip, keys, values, end_bit=0):
    n = keys.count
    kb = hip.DeviceArray(n, np.uint64)
    vb = hip.DeviceArray(n, np.uint32)
    f = hip._lib.hipcstone_sort_by_key_u64_noalloc
    st = f(None, C.c_void_p(keys.ptr), C.c_void_p(values.ptr), C.c_size_t(n), C.c_void_p(kb.ptr), C.c_void_p(vb.ptr),
           C.c_int(end_bit))
    hip._check(st, "sort_by_key_u64_noalloc")
    hip.device_sync()


def _case(hip, n, seed, end_bit):
    rng = np.random.default_rng(seed)
    # duplicate keys (a narrow key range) so stability is exercised, not only ordering
    keys = rng.integers(0, 1 << 20, n, dtype=np.uint64) << np.uint64(40)
    keys |= rng.integers(0, 1 << 40, n, dtype=np.uint64) & np.uint64((1 << 40) - 1 if end_bit == 0 else 0)
    vals = rng.permutation(n).astype(np.uint32)
    a_k, a_v = hip.DeviceArray.from_numpy(keys), hip.DeviceArray.from_numpy(vals)
    b_k, b_v = hip.DeviceArray.from_numpy(keys), hip.DeviceArray.from_numpy(vals)
    print(f"no-alloc sort: n={n} seed={seed} end_bit={end_bit}", flush=True)
    _sort_noalloc(hip, a_k, a_v, end_bit)
    hip.sort_by_key(b_k, b_v, end_bit=end_bit)
    ak, av = a_k.download(), a_v.download()
    bk, bv = b_k.download(), b_v.download()
    assert np.array_equal(ak, bk) and np.array_equal(av, bv)
    # and both are the stable sort of the input
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ak, keys[order]) and np.array_equal(av, vals[order])


def test_noalloc_sort_matches_workspace_sort_twice(hip):
    _case(hip, 3_000_001, 11, 0)  # odd n, full 64-bit keys (8 passes)
    _case(hip, 200_003, 12, 0)    # smaller: the scratch holds the first call's look-back words
    _case(hip, 1 << 24, 13, 64)   # the maximum size
    _case(hip, 777, 14, 0)        # one tile


def test_noalloc_sort_rejects_oversize(hip):
    n = (1 << 24) + 1
    f = hip._lib.hipcstone_sort_by_key_u64_noalloc
    # sizes are checked before any pointer is used
    st = f(None, C.c_void_p(0), C.c_void_p(0), C.c_size_t(n), C.c_void_p(0), C.c_void_p(0), C.c_int(0))
    assert st == 1  # HIPCSTONE_INVALID_ARG
    assert b"2^24" in hip._lib.hipcstone_last_error()
