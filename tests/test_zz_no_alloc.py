"""sph_sort_pairs_noalloc: the sort for callers that cannot allocate, checked
through the C ABI against sph_sort_pairs (the allocating entry point it
varies) and against numpy's stable argsort, bit for bit, on two calls in a
row (the second smaller, so it runs on scratch the first left behind), and
its size limit."""
import ctypes as C

import numpy as np
import pytest
import torch

from fluids_amd._lib import lib

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 20
SPH_ERR_RANGE = -4


def _sort(fn, keys, values, key_bits):
    k = torch.from_numpy(keys.astype(np.uint32).view(np.int32)).to("cuda")
    v = torch.from_numpy(values.astype(np.int32)).to("cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = getattr(lib(), fn)(C.c_void_p(k.data_ptr()), C.c_void_p(v.data_ptr()), len(keys), key_bits,
                            C.c_void_p(stream))
    torch.cuda.synchronize()
    assert rc == 0, lib().sph_last_error().decode()
    return k.cpu().numpy().view(np.uint32), v.cpu().numpy()


@pytest.mark.parametrize("key_bits", [27, 0])
def test_noalloc_sort_equals_allocating_sort_twice(key_bits):
    rng = np.random.default_rng(key_bits + 5)
    for n in (NOALLOC_MAX, 300_001):  # the second call smaller than the first
        print("case noalloc sort", key_bits, n, flush=True)
        keys = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 2] = np.sort(keys[: n // 2] & np.uint32(0x7FFFFFF))  # a nearly sorted half
        values = np.arange(n, dtype=np.int32)[::-1].copy()
        ak, av = _sort("sph_sort_pairs", keys, values, key_bits)
        nk, nv = _sort("sph_sort_pairs_noalloc", keys, values, key_bits)
        assert np.array_equal(nk, ak) and np.array_equal(nv, av)
        mask = np.uint32(0xFFFFFFFF) if key_bits == 0 else np.uint32((1 << key_bits) - 1)
        order = np.argsort(keys & mask, kind="stable")
        assert np.array_equal(nk, keys[order]) and np.array_equal(nv, values[order])


def test_noalloc_sort_rejects_sizes_past_its_scratch():
    k = torch.zeros(16, dtype=torch.int32, device="cuda")
    v = torch.zeros(16, dtype=torch.int32, device="cuda")
    rc = lib().sph_sort_pairs_noalloc(C.c_void_p(k.data_ptr()), C.c_void_p(v.data_ptr()), NOALLOC_MAX + 1, 0,
                                      None)
    assert rc == SPH_ERR_RANGE
    assert "SPH_SORT_NOALLOC_MAX" in lib().sph_last_error().decode()
    assert lib().sph_sort_pairs_noalloc(C.c_void_p(k.data_ptr()), C.c_void_p(v.data_ptr()), 0, 0, None) == 0
