"""The sort of a nearly sorted step (sph_sort_near, and the single-scene
solvers' per-step sort built on the same kernels) against numpy's stable
argsort: for every case the sorted keys equal keys[order] and the
permutation equals order = argsort(keys, kind="stable"), bit for bit. The
cases cover the kernels' paths: elements that keep their key, moved elements
alone at their insertion point, groups of 2-64 (ranked per element), of
65-4096 (sorted in LDS) and beyond (sorted in place), ties between moved and
unmoved equal keys, hints unrelated to the keys, the sizes around the
64-element mask words and 1024-position scan tiles, and a size past 1024
tiles (the tile sums then go through the separate scan launch). The solver test runs a
scene with SPH_FULL_SORT (the radix sort of every key each step) beside the
default and compares every field after each step."""
import ctypes as C

import numpy as np
import pytest
import torch

from fluids_amd._lib import lib, SPH_OK

pytestmark = pytest.mark.gpu

SPH_ERR_INVALID = -1
SPH_FULL_SORT = 4


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to("cuda")


def sort_near(hint, keys, in_place=False):
    h, k = _dev(hint.astype(np.uint32)), _dev(keys.astype(np.uint32))
    n = len(keys)
    out = h if in_place else torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    perm = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib().sph_sort_near(C.c_void_p(h.data_ptr()), C.c_void_p(k.data_ptr()), C.c_void_p(out.data_ptr()),
                             C.c_void_p(perm.data_ptr()), n, C.c_void_p(stream))
    torch.cuda.synchronize()
    assert rc == SPH_OK, lib().sph_last_error().decode()
    return out.cpu().numpy().view(np.uint32)[:n], perm.cpu().numpy()[:n]


def check(hint, keys, in_place=False):
    gk, gp = sort_near(hint, keys, in_place)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(gp, order), f"first difference at {np.argmax(gp != order)}"
    assert np.array_equal(gk, keys[order])


def near_sorted(rng, n, frac, span, bits=27):
    hint = np.sort(rng.integers(0, 2 ** bits, n, dtype=np.uint64)).astype(np.int64)
    keys = hint.copy()
    sel = rng.random(n) < frac
    keys[sel] = np.clip(keys[sel] + rng.integers(-span, span + 1, int(sel.sum())), 0, 2 ** bits - 1)
    return hint.astype(np.uint32), keys.astype(np.uint32)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 4095, 4096, 4097, 8191, 300_000, 1_000_000, 1_500_000])
def test_moved_fraction_with_small_moves(n):
    print("case near-sorted", n, flush=True)
    rng = np.random.default_rng(n)
    for frac, span in [(0.0, 1), (0.05, 64), (0.3, 5000), (1.0, 2 ** 20)]:
        check(*near_sorted(rng, n, frac, span))


@pytest.mark.parametrize("n", [100, 5000, 250_000])
def test_hint_unrelated_to_keys(n):
    print("case unrelated hint", n, flush=True)
    rng = np.random.default_rng(n + 1)
    hint = np.sort(rng.integers(0, 2 ** 30, n, dtype=np.uint64)).astype(np.uint32)
    keys = rng.integers(0, 2 ** 30, n, dtype=np.uint64).astype(np.uint32)
    check(hint, keys)
    keys = rng.integers(0, 16, n, dtype=np.uint64).astype(np.uint32)  # many ties, few values
    check(hint, keys)


@pytest.mark.parametrize("n,size", [(3000, 3000), (4096, 4096), (4097, 4097), (200_000, 150_000),
                                    (10_000, 65), (10_000, 64)])
def test_one_group(n, size):
    """`size` elements move to one key that lies in a gap of the hint (the
    non-finite-position case: every such particle hashes to one cell)."""
    print("case one group", n, size, flush=True)
    rng = np.random.default_rng(size)
    hint = np.sort(rng.integers(100, 2 ** 30, n, dtype=np.uint64)).astype(np.uint32)
    keys = hint.copy()
    sel = rng.choice(n, size, replace=False)
    keys[sel] = 7
    check(hint, keys)
    keys[sel] = rng.integers(0, 50, size, dtype=np.uint64).astype(np.uint32)  # one gap, unequal keys
    check(hint, keys)


def test_many_groups_of_every_size():
    print("case many groups", flush=True)
    rng = np.random.default_rng(5)
    n = 400_000
    hint = (np.sort(rng.integers(0, 2 ** 20, n, dtype=np.uint64)) * 1024).astype(np.uint32)  # gaps of >= 1023
    keys = hint.copy()
    for size in [2, 3, 17, 63, 64, 65, 100, 1000, 4095, 4096, 4097, 9000]:
        idx = rng.choice(n, size, replace=False)
        gap = int(hint[rng.integers(0, n)])
        keys[idx] = gap + 1 + rng.integers(0, 1000, size).astype(np.uint32)
    check(hint, keys)


def test_ties_between_moved_and_unmoved():
    print("case ties", flush=True)
    rng = np.random.default_rng(9)
    n = 100_000
    hint = np.sort(rng.integers(0, 500, n, dtype=np.uint64)).astype(np.uint32)
    keys = hint.copy()
    sel = rng.random(n) < 0.2
    keys[sel] = rng.integers(0, 500, int(sel.sum()), dtype=np.uint64).astype(np.uint32)
    check(hint, keys)


def test_in_place_output_and_repeated_calls():
    """sorted_keys may be the hint; the work arrays start from zero on every
    call, and the library's sort stays exact when chained step after step."""
    print("case in place", flush=True)
    rng = np.random.default_rng(11)
    hint, keys = near_sorted(rng, 70_000, 0.1, 300)
    for _ in range(3):
        check(hint, keys, in_place=True)
        hint = np.sort(keys)
        keys = hint.copy()
        sel = rng.random(len(keys)) < 0.1
        keys[sel] = keys[sel] + rng.integers(0, 300, int(sel.sum())).astype(np.uint32)


def test_bad_arguments():
    L = lib()
    assert L.sph_sort_near(None, None, None, None, -1, None) == SPH_ERR_INVALID
    assert L.sph_sort_near(None, None, None, None, 10, None) == SPH_ERR_INVALID
    assert L.sph_sort_near(None, None, None, None, 0, None) == SPH_OK
    h = _dev(np.array([1, 2, 3, 0], np.uint32))  # decreases
    k = _dev(np.array([1, 2, 3, 4], np.uint32))
    o = torch.zeros(4, dtype=torch.int32, device="cuda")
    p = torch.zeros(4, dtype=torch.int32, device="cuda")
    ptr = lambda t: C.c_void_p(t.data_ptr())
    assert L.sph_sort_near(ptr(h), ptr(k), ptr(o), ptr(p), 4, None) == SPH_ERR_INVALID
    assert b"non-decreasing" in L.sph_last_error()
    h2 = _dev(np.array([0, 1, 2, 3], np.uint32))
    assert L.sph_sort_near(ptr(h2), ptr(k), ptr(k), ptr(p), 4, None) == SPH_ERR_INVALID
    torch.cuda.synchronize()


@pytest.mark.parametrize("method", ["pcisph", "wcsph"])
def test_solver_sort_paths_give_identical_steps(method):
    """A scene stepped with the moved-keys sort and with the full radix sort
    (SPH_FULL_SORT): keys, permutation, lists and the state fields equal
    after each of 30 steps."""
    import dataclasses
    from fluids_amd import scenes
    print("case solver paths", method, flush=True)
    base = scenes.c2_pcisph_double_dambreak(spacing=0.02) if method == "pcisph" else scenes.small_wcsph(seed=3)
    sims = []
    for flags in (0, SPH_FULL_SORT):
        st = dict(base.settings)
        st["flags"] = st.get("flags", 0) | flags
        sims.append(scenes.build(dataclasses.replace(base, settings=st)))
    fields = ["keys", "perm", "ncount", "nlist", "bncount", "pos", "vel", "density", "pressure"]
    for step in range(30):
        for s in sims:
            s.step()
        for f in fields:
            a, b = sims[0].field(f), sims[1].field(f)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (step, f)


def test_solver_sort_paths_agree_when_many_positions_turn_non_finite():
    """Failure path forced from the host: 40 % of a scene's velocities are
    set to NaN, so after one step those particles (and, through viscosity,
    their neighbours) have non-finite positions, which all hash to one cell:
    the next step's sort sees one group of several hundred moved particles
    (sorted by a workgroup). The moved-keys path and the full radix sort give
    the same bits in every field for three steps."""
    import dataclasses
    from fluids_amd import scenes
    base = scenes.small_pcisph(seed=5)
    print("case nan group", base.n_fluid, flush=True)
    sims = []
    for flags in (0, SPH_FULL_SORT):
        st = dict(base.settings)
        st["flags"] = st.get("flags", 0) | flags
        sims.append(scenes.build(dataclasses.replace(base, settings=st)))
    rng = np.random.default_rng(1)
    vel = sims[0].velocities()
    bad = rng.random(vel.shape[0]) < 0.4
    vel[bad, :3] = np.nan
    for s in sims:
        s.set_velocities(vel)
    fields = ["keys", "perm", "ncount", "nlist", "pos", "vel", "density", "pressure"]
    for step in range(3):
        for s in sims:
            s.step()
        for f in fields:
            a, b = sims[0].field(f), sims[1].field(f)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (step, f)
    nan = int((~np.isfinite(sims[0].field("pos")[:, :3]).all(1)).sum())
    assert nan >= int(bad.sum())
    assert sims[0].step_stats()["nan_positions"] == sims[1].step_stats()["nan_positions"] == nan


def test_solver_sort_paths_agree_over_400_steps_at_1m():
    """The 1M throughput scene (PCISPH, tank walls) stepped 400 times with
    the moved-keys sort and with the full radix sort: the sorted keys, the
    permutation, the lists and the state are equal after every 100 steps
    (late in the run 8-12 % of the keys change cell every step)."""
    import dataclasses
    from fluids_amd import scenes
    base = scenes.c3_pcisph_throughput(1_000_000)
    print("case 1M sort paths", base.n_fluid, base.n_boundary, flush=True)
    sims = []
    for flags in (0, SPH_FULL_SORT):
        st = dict(base.settings)
        st["flags"] = st.get("flags", 0) | flags
        sims.append(scenes.build(dataclasses.replace(base, settings=st)))
    fields = ["keys", "perm", "ncount", "nlist", "bncount", "pos", "vel", "density"]
    for chunk in range(4):
        for s in sims:
            s.step(100)
        for f in fields:
            a, b = sims[0].field(f), sims[1].field(f)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (100 * (chunk + 1), f)
        print("equal after", 100 * (chunk + 1), "steps", flush=True)
    for s in sims:
        s.close()
