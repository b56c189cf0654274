"""The PCISPH throughput scenes at their full sizes.

* 1M particles (the measured workload): two steps on the HIP path and on the
  CPU oracle from the same state, every integer field and every fp32 field
  compared bit for bit, as in the small parity cases; and one step from the
  state after 300 GPU steps (collapsed boxes, fluid on the floor and walls).
* 4M particles: the oracle is not run; the GPU's outputs are checked through
  properties that fix them independently of any implementation:
    - the sorted keys are the Morton keys of the sorted positions' cells,
      non-decreasing, and the permutation is one;
    - for a seeded sample of particles, the neighbour list equals a numpy
      restatement of compute_neigh_list (grid.cl:145-181): own cell first,
      then the 26 neighbours in lexicographic (dx, dy, dz) order, sorted
      order inside a cell, strict r^2 < h^2 in fp32, cap 50;
    - for sampled particles without boundary neighbours, the density is
      W * ((sum over the list of (H^2 - r^2)^3) * m) accumulated in list
      order in fp32 (pcisph.cl:33-102).
"""
import numpy as np
import pytest

import oracle
from fluids_amd import scenes
from helpers import build_oracle, check_grid_properties, max_rel

pytestmark = pytest.mark.gpu

INT_FIELDS = ["keys", "perm", "ncount", "nlist", "bncount", "bnlist"]
PCI_FIELDS = ["pos_sorted", "vel_sorted", "density", "normal", "force_other", "pressure", "err",
              "force_pressure", "pos_pred", "pos", "vel"]


def test_pcisph_1m_steps_bitexact():
    sc = scenes.c3_pcisph_throughput(1_000_000)
    print("case pcisph 1M", sc.n_fluid, sc.n_boundary, flush=True)
    sim = scenes.build(sc)
    orc = build_oracle(sc)
    assert np.array_equal(sim.positions(), orc.field("pos")), "relaxed initial state differs"
    for step in range(2):
        orc.set_state(sim.positions(), sim.velocities())
        sim.step(1)
        orc.step(1)
        assert sim.step_stats()["iterations"] == orc.iterations
        for f in INT_FIELDS:
            a, b = sim.field(f), orc.field(f)
            assert np.array_equal(a, b), f"step {step}: {f} differs at {np.argwhere(a != b)[:5].ravel()}"
        for f in PCI_FIELDS:
            a, b = sim.field(f), orc.field(f)
            if a.ndim == 2:
                a, b = a[:, :3], b[:, :3]
            nbad = int(np.sum(a.view(np.uint32) != b.view(np.uint32)))
            assert nbad == 0, f"step {step}: {f} has {nbad} differing words (max-rel {max_rel(a, b):.2e})"
        print("1M step", step, "bit-exact; mean list length", float(sim.field("ncount").mean()), flush=True)


def test_pcisph_1m_late_step_bitexact():
    """The same scene after 300 GPU steps, when the boxes have collapsed and
    spread over the floor (longer lists, about one boundary neighbour per
    particle): one step from the GPU's state on both sides, bit for bit."""
    sc = scenes.c3_pcisph_throughput(1_000_000)
    print("case pcisph 1M late", sc.n_fluid, sc.n_boundary, flush=True)
    sim = scenes.build(sc)
    orc = build_oracle(sc)
    sim.step(300)
    orc.set_state(sim.positions(), sim.velocities())
    sim.step(1)
    orc.step(1)
    st = sim.step_stats()
    assert st["iterations"] == orc.iterations
    assert st["boundary_neighbours"] > 300_000, st  # the fluid reaches the walls
    assert st["fluid_neighbours"] == int(orc.field("ncount").astype(np.int64).sum())
    assert st["boundary_neighbours"] == int(orc.field("bncount").astype(np.int64).sum())
    for f in INT_FIELDS:
        a, b = sim.field(f), orc.field(f)
        assert np.array_equal(a, b), f"late: {f} differs at {np.argwhere(a != b)[:5].ravel()}"
    for f in PCI_FIELDS:
        a, b = sim.field(f), orc.field(f)
        if a.ndim == 2:
            a, b = a[:, :3], b[:, :3]
        nbad = int(np.sum(a.view(np.uint32) != b.view(np.uint32)))
        assert nbad == 0, f"late: {f} has {nbad} differing words (max-rel {max_rel(a, b):.2e})"
    print("1M late step bit-exact; mean list lengths", st["fluid_neighbours"] / sc.n_fluid,
          st["boundary_neighbours"] / sc.n_fluid, flush=True)


def test_pcisph_4m_properties():
    sc = scenes.c3_pcisph_throughput(4_000_000)
    print("case pcisph 4M", sc.n_fluid, sc.n_boundary, flush=True)
    sim = scenes.build(sc)
    sim.step(2)
    assert sim.particle_count() == 4_000_000
    consts = oracle.Oracle(sc.settings).constants()
    lists, dens = check_grid_properties(sim.field, consts, sc.settings["fluid_support_radius"], 2_000, 404)
    assert dens > 1_000
    assert np.all(np.isfinite(sim.positions()))
    print("4M: sample of", lists, "lists equal the restatement;", dens, "densities exact", flush=True)
