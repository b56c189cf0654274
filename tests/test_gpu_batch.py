"""Batched WCSPH (sph_batch_*) against the same scenes run one at a time.

A scene advanced inside a batch evaluates the same per-particle functions on
the same inputs as the single-scene solver, so its stored state must be equal
bit for bit after every chunk of steps; the chunks are replays of the step's
HIP graph with host reads between them. The autotuner's choice must not
change any result, and its cache must round-trip.
"""
import numpy as np
import pytest

from fluids_amd import scenes

pytestmark = pytest.mark.gpu


def _small_batch(count=24):
    return scenes.c4_batched(seed=404, count=count, n_min=1000, n_max=6000)


def test_batch_bitexact_vs_single_scenes():
    scs = _small_batch()
    b = scenes.build_batch(scs, autotune=True, trials=1)
    singles = [scenes.build(sc) for sc in scs]
    print("launch shapes:", {k: b.launch(k) for k in range(5)}, flush=True)
    for chunk in range(4):
        b.step(10)
        for s in singles:
            s.step(10)
        for k, (sc, s) in enumerate(zip(scs, singles)):
            label = f"chunk {chunk} scene {k} (n={sc.n_fluid})"
            assert np.array_equal(b.positions(k), s.positions()), label
            assert np.array_equal(b.velocities(k), s.velocities()), label
            assert np.array_equal(b.half_velocities(k), s.half_velocities()), label
    for k, s in enumerate(singles):
        assert np.array_equal(b.density(k), s.field("density")), k
        assert np.array_equal(b.neighbour_counts(k), s.field("ncount")), k
    assert b.total_particles == sum(sc.n_fluid for sc in scs)


def test_graph_replay_matches_eager_steps():
    scs = _small_batch(8)
    g = scenes.build_batch(scs, autotune=False)
    e = scenes.build_batch(scs, autotune=False)
    e.set_graph(False)
    for chunk in range(3):
        g.step(10)
        e.step(10)
        for k in range(len(scs)):
            assert np.array_equal(g.positions(k), e.positions(k)), (chunk, k)
            assert np.array_equal(g.half_velocities(k), e.half_velocities(k)), (chunk, k)


def test_launch_shapes_do_not_change_results():
    scs = _small_batch(6)
    a = scenes.build_batch(scs, autotune=False)
    b = scenes.build_batch(scs, autotune=False)
    b.set_launch("batch_lists_density", 64, 512, 0)
    b.set_launch("batch_normals", 128, 1024, 24)
    b.set_launch("batch_accel_integrate", 64, 256, 24)
    b.set_launch("batch_hash", 128, 64, 24)
    b.set_launch("batch_gather_ranges", 64, 128, 24)
    a.step(20)
    b.step(20)
    for k in range(len(scs)):
        assert np.array_equal(a.positions(k), b.positions(k)), k


def test_autotune_cache_round_trip(tmp_path):
    scs = _small_batch(6)
    cache = str(tmp_path / "tune.txt")
    a = scenes.build_batch(scs, autotune=False)
    assert a.autotune(cache, 1) is False
    shapes = [a.launch(k) for k in range(5)]
    b = scenes.build_batch(scs, autotune=False)
    assert b.autotune(cache, 1) is True
    assert [b.launch(k) for k in range(5)] == shapes
    # tuning restored the state: both batches continue identically
    a.step(5)
    b.step(5)
    for k in range(len(scs)):
        assert np.array_equal(a.positions(k), b.positions(k)), k


def test_c4_full_batch_against_oracle_and_single_scenes():
    """The configured batch: 512 seeded scenes of 1k-64k particles (log-uniform,
    fluids_amd.scenes.c4_batched), 20 graph-replayed steps. Eight scenes drawn
    from a seed are stepped by the oracle from the same initial state and must
    be equal bit for bit (positions, velocities, half-step velocities, and the
    last step's densities and neighbour counts); all 512 must equal the same
    scenes run as separate single-scene solvers."""
    import os

    import oracle
    from helpers import build_oracle
    oracle.set_threads(min(16, os.cpu_count() or 1))
    scs = scenes.c4_batched()
    assert len(scs) == 512
    sizes = np.array([sc.n_fluid for sc in scs])
    print("case c4 full batch", len(scs), int(sizes.sum()), int(sizes.min()), int(sizes.max()), flush=True)
    b = scenes.build_batch(scs, autotune=False)
    steps = 20
    b.step(steps)
    b.synchronize()
    pick = np.random.default_rng(404).choice(len(scs), 8, replace=False)
    pick[0] = int(np.argmax(sizes))  # the largest scene is always among them
    for k in pick:
        orc = build_oracle(scs[k])
        orc.step(steps)
        label = f"scene {k} (n={scs[k].n_fluid})"
        assert np.array_equal(b.positions(k), orc.field("pos")), label
        assert np.array_equal(b.velocities(k), orc.field("vel")), label
        assert np.array_equal(b.half_velocities(k), orc.field("vel_half")), label
        assert np.array_equal(b.density(k), orc.field("density")), label
        assert np.array_equal(b.neighbour_counts(k), orc.field("ncount")), label
        print("oracle equal:", label, flush=True)
    for k, sc in enumerate(scs):
        s = scenes.build(sc)
        s.step(steps)
        assert np.array_equal(b.positions(k), s.positions()), k
        assert np.array_equal(b.velocities(k), s.velocities()), k
        assert np.array_equal(b.half_velocities(k), s.half_velocities()), k
        s.close()
    b.close()


def test_tiny_empty_and_clumped_scenes_in_a_batch():
    """Scenes at the edges of the batched sort and walk, between ordinary
    scenes: one particle, three, none, and 120 particles packed into a third
    of a cell (lists at the cap; the clump flies apart, so most of its keys
    change cell every step). Each non-empty scene must equal its single-scene
    run bit for bit after every chunk of graph-replayed steps."""
    base = _small_batch(4)
    rng = np.random.default_rng(7)

    def variant(sc, parts, name):
        vols = [parts.astype(np.float32)] if len(parts) else []
        return scenes.Scene(name, dict(sc.settings), sc.container, vols)

    h = base[2].settings["fluid_support_radius"]
    c = base[2].volumes[0][0, :3].astype(np.float64)
    clump = np.concatenate([c + rng.uniform(0.0, h / 3, (120, 3)), np.ones((120, 1))], axis=1)
    scs = [base[0], variant(base[0], base[0].volumes[0][:1], "one"),
           variant(base[1], base[1].volumes[0][:3], "three"), variant(base[3], np.zeros((0, 4)), "empty"),
           variant(base[2], clump, "clump"), base[3]]
    b = scenes.build_batch(scs, autotune=False)
    singles = {k: scenes.build(sc) for k, sc in enumerate(scs) if sc.n_fluid > 0}
    assert b.total_particles == sum(sc.n_fluid for sc in scs)
    for chunk, steps in enumerate([1, 10, 10, 9]):
        b.step(steps)
        for s in singles.values():
            s.step(steps)
        for k, s in singles.items():
            label = f"chunk {chunk} scene {scs[k].name} (n={scs[k].n_fluid})"
            assert np.array_equal(b.positions(k), s.positions()), label
            assert np.array_equal(b.velocities(k), s.velocities()), label
            assert np.array_equal(b.half_velocities(k), s.half_velocities()), label
            assert np.array_equal(b.density(k), s.field("density")), label
            assert np.array_equal(b.neighbour_counts(k), s.field("ncount")), label
        if chunk == 0:
            assert int(b.neighbour_counts(4).max()) == 50  # the clump's lists reach the cap
    for s in singles.values():
        s.close()
    b.close()


def test_batch_moved_keys_sort_matches_full_sort_singles_over_200_steps():
    """The batch sorts only the keys that changed cell after its first step;
    the same scenes run one at a time with SPH_FULL_SORT (the radix sort of
    every key every step) must give the same bits after every 50 of 200
    graph-replayed steps."""
    import dataclasses
    scs = _small_batch(24)
    b = scenes.build_batch(scs, autotune=False)
    singles = []
    for sc in scs:
        st = dict(sc.settings)
        st["flags"] = st.get("flags", 0) | 4  # SPH_FULL_SORT
        singles.append(scenes.build(dataclasses.replace(sc, settings=st)))
    for chunk in range(4):
        b.step(50)
        for s in singles:
            s.step(50)
        for k, s in enumerate(singles):
            label = f"step {50 * (chunk + 1)} scene {k} (n={scs[k].n_fluid})"
            assert np.array_equal(b.positions(k), s.positions()), label
            assert np.array_equal(b.velocities(k), s.velocities()), label
            assert np.array_equal(b.half_velocities(k), s.half_velocities()), label
    for k, s in enumerate(singles):
        assert np.array_equal(b.density(k), s.field("density")), k
        assert np.array_equal(b.neighbour_counts(k), s.field("ncount")), k
        s.close()
    b.close()


def test_mixed_density_forms_in_one_batch():
    """Half the scenes keep SPH_FIX_WCSPH_DENSITY (W_poly6 evaluated on r^2),
    half use the reference's own form (W_poly6 evaluated on r,
    wcsph.cl:61-62,70-71); the batched walk picks the form per scene, and
    every scene must equal its single-scene run."""
    import dataclasses
    base = _small_batch(8)
    scs = []
    for k, sc in enumerate(base):
        st = dict(sc.settings)
        if k % 2:
            st["flags"] = st.get("flags", 0) & ~1  # SPH_FIX_WCSPH_DENSITY off
        scs.append(dataclasses.replace(sc, settings=st))
    assert {sc.settings["flags"] & 1 for sc in scs} == {0, 1}
    b = scenes.build_batch(scs, autotune=False)
    singles = [scenes.build(sc) for sc in scs]
    for chunk in range(3):
        b.step(10)
        for s in singles:
            s.step(10)
        for k, s in enumerate(singles):
            label = f"chunk {chunk} scene {k} flags {scs[k].settings['flags']}"
            assert np.array_equal(b.positions(k), s.positions()), label
            assert np.array_equal(b.velocities(k), s.velocities()), label
            assert np.array_equal(b.density(k), s.field("density")), label
    for s in singles:
        s.close()
    b.close()
