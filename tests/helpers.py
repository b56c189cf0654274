"""Shared test helpers: oracle construction and field comparison."""
import numpy as np

import oracle


def build_oracle(sc, twin=False):
    """The scene on the CPU oracle, in the reference's init order."""
    o = oracle.Oracle(sc.settings, twin=twin)
    o.set_rect_limits(*sc.container)
    for v in sc.volumes:
        o.add_volume(v)
    for b in sc.bodies:
        o.add_boundary(b.samples, b.can_move, b.mass, b.R, b.com)
    return o


def max_rel(a, b):
    """max |a - b| / max |b|: the field-wise max-relative error."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


# ----------------------------------------------------------------------------
# An independent restatement of the grid path's integer outputs, for checks at
# sizes where a second full implementation is too slow: Morton keys of cells
# (kernels/morton.h:37-56, kernels/grid.h:74-91) and compute_neigh_list
# (grid.cl:145-181) for sampled particles, plus the PCISPH initial density
# (pcisph.cl:33-102) of sampled particles without boundary neighbours.
f32 = np.float32


def split_by_3(a):
    x = a.astype(np.uint32) & np.uint32(0xFFFF)
    for sh, m in ((16, 0xFF0000FF), (8, 0x0F00F00F), (4, 0xC30C30C3), (2, 0x49249249)):
        x = (x | (x << np.uint32(sh))) & np.uint32(m)
    return x


def morton(c):
    return split_by_3(c[:, 0]) | (split_by_3(c[:, 1]) << np.uint32(1)) | (split_by_3(c[:, 2]) << np.uint32(2))


def cells_of(p, cell, cps):
    c = np.floor((p + f32(2.5)) / f32(cell)).astype(np.int64)
    return (c + cps) % cps


NEIGH_OFFSETS = [(0, 0, 0)] + [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)
                               if (dx, dy, dz) != (0, 0, 0)]


def check_grid_properties(get, consts, cell, n_sample, seed):
    """get(name) returns a field of a solver after a PCISPH step. Returns the
    number of sampled lists and densities that were compared."""
    cps = int(consts["grid_cps"])
    h2 = f32(consts["grid_sqr_support"])
    keys, perm, ps = get("keys"), get("perm"), get("pos_sorted")[:, :3]
    n = keys.shape[0]
    assert np.all(keys[1:] >= keys[:-1]), "keys not sorted"
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.int32)), "perm is not a permutation"
    c = cells_of(ps, cell, cps)
    assert np.array_equal(morton(c), keys), "keys are not the Morton keys of the sorted cells"
    nc, nl, bnc, dens = get("ncount"), get("nlist").reshape(-1, n), get("bncount"), get("density")
    sample = np.random.default_rng(seed).choice(n, n_sample, replace=False)
    H2 = f32(consts["pci_h_d"]) * f32(consts["pci_h_d"])
    M, W = f32(consts["pci_mass_d"]), f32(consts["pci_default_eval"])
    densities = 0
    for i in sample:
        pi = ps[i]
        expect = []
        for d3 in NEIGH_OFFSETS:
            q = c[i] + np.array(d3)
            if np.any(q < 0) or np.any(q >= cps):
                continue  # the scenes checked here lie far from the grid's faces
            k = morton(q[None, :])[0]
            a, b = np.searchsorted(keys, k, "left"), np.searchsorted(keys, k, "right")
            for j in range(a, b):
                r = pi - ps[j]
                r2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]
                if r2 < h2:
                    expect.append((j, r2))
                    if len(expect) == 50:
                        break
            if len(expect) == 50:
                break
        got = nl[: nc[i], i]
        assert nc[i] == len(expect) and np.array_equal(got, [j for j, _ in expect]), f"list of particle {i}"
        if bnc[i] == 0:
            d = f32(0)
            for _, r2 in expect:
                t = H2 - r2
                d = d + (t * t) * t
            d = W * ((d * M) + f32(0))
            assert d == dens[i], (i, d, dens[i])
            densities += 1
    return len(sample), densities
