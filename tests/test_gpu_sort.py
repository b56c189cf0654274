"""sph_sort_pairs, the library's stable key/value sort on caller-owned device
arrays (the reference's clsort, opencl/algorithms/clsort.h:12-35), against
numpy's stable argsort: for every size and key width the sorted keys and the
carried values equal keys[order] and values[order] with
order = argsort(keys & mask, kind="stable"), bit for bit. Device memory comes
from torch; the sort runs through the C ABI."""
import ctypes as C

import numpy as np
import pytest
import torch

from fluids_amd._lib import lib

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to("cuda")


def sort_on_gpu(keys, values, key_bits, fn="sph_sort_pairs"):
    k, v = _dev(keys.astype(np.uint32)), _dev(values.astype(np.int32))
    stream = torch.cuda.current_stream().cuda_stream
    rc = getattr(lib(), fn)(C.c_void_p(k.data_ptr()), C.c_void_p(v.data_ptr()), len(keys), key_bits,
                            C.c_void_p(stream))
    torch.cuda.synchronize()
    assert rc == 0, lib().sph_last_error().decode()
    return k.cpu().numpy().view(np.uint32), v.cpu().numpy()


def expected(keys, values, key_bits):
    mask = np.uint32(0xFFFFFFFF) if key_bits in (0, 32) else np.uint32((1 << key_bits) - 1)
    order = np.argsort(keys & mask, kind="stable")
    return keys[order], values[order]


@pytest.mark.parametrize("n", [0, 1, 63, 256, 2047, 2048, 2049, 65536, 65537, 1_000_000])
@pytest.mark.parametrize("key_bits", [8, 27, 0])
def test_sort_pairs_equals_stable_argsort(n, key_bits):
    print("case sort", n, key_bits, flush=True)
    rng = np.random.default_rng(n * 7 + key_bits)
    keys = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    if key_bits == 8:
        keys &= np.uint32(0x3FF)  # many equal digits: stability shows
    values = rng.integers(-2 ** 31, 2 ** 31, n, dtype=np.int64).astype(np.int32)
    gk, gv = sort_on_gpu(keys, values, key_bits)
    ek, ev = expected(keys, values, key_bits)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


@pytest.mark.parametrize("key_bits", [9, 16, 21, 30, 32])
def test_sort_pairs_nearly_sorted_morton_keys(key_bits):
    """The solver's case: keys sorted but for a few percent moved elsewhere,
    values the identity (the permutation the step gathers with)."""
    rng = np.random.default_rng(key_bits)
    n = 300_000
    keys = np.sort(rng.integers(0, 2 ** key_bits, n, dtype=np.uint64)).astype(np.uint32)
    movers = rng.choice(n, n // 20, replace=False)
    keys[movers] = rng.integers(0, 2 ** key_bits, len(movers), dtype=np.uint64).astype(np.uint32)
    values = np.arange(n, dtype=np.int32)
    gk, gv = sort_on_gpu(keys, values, key_bits)
    ek, ev = expected(keys, values, key_bits)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


def test_sort_pairs_rejects_bad_arguments():
    L = lib()
    assert L.sph_sort_pairs(None, None, -1, 0, None) == -1
    assert L.sph_sort_pairs(None, None, 10, 33, None) == -1
    assert L.sph_sort_pairs(None, None, 10, 0, None) == -1
    assert L.sph_sort_pairs(None, None, 0, 0, None) == 0
