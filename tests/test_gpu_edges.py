"""Edge cases of the HIP path against the CPU oracle, bit for bit:

* a dense clump (three interleaved lattices) where most lists reach the
  cap of 50 (NEIGH_LIST_MAX_LENGTH, kernels/common.h:4; the walk stops at the
  cap, grid.cl:170-177), for PCISPH and WCSPH, with the overflow and
  edge-read counters;
* ragged particle counts around the wave (64) and workgroup (256) sizes,
  down to a single particle;
* a step with no fluid particles at all (only the tank);
* the correction loop's host chunks (a relaxation step of 10000
  corrections);
* failure detection forced from the host: a NaN velocity (counted NaN
  positions) and particles set outside the container (the range box grows;
  edge reads counted as the oracle counts them).
"""
import numpy as np
import pytest

from fluids_amd import scenes
from fluids_amd.scenes import Rng, Scene
from helpers import build_oracle

pytestmark = pytest.mark.gpu

INT_FIELDS = ["keys", "perm", "ncount", "nlist", "bncount", "bnlist"]


def _step_and_compare(sc, method, steps, fields, label):
    sim = scenes.build(sc)
    orc = build_oracle(sc)
    for step in range(steps):
        pos, vel = sim.positions(), sim.velocities()
        vh = sim.half_velocities() if method == 0 else None
        orc.set_state(pos, vel, vh)
        before = orc.edge_reads
        sim.step(1)
        orc.step(1)
        for f in INT_FIELDS:
            a, b = sim.field(f), orc.field(f)
            assert np.array_equal(a, b), f"{label} step {step}: {f} differs"
        for f in fields:
            a, b = sim.field(f), orc.field(f)
            if a.ndim == 2:
                a, b = a[:, :3], b[:, :3]
            nbad = int(np.sum(a.view(np.uint32) != b.view(np.uint32)))
            assert nbad == 0, f"{label} step {step}: {f} has {nbad} differing words"
        st = sim.step_stats()
        assert st["fluid_list_overflows"] == int(np.sum(orc.field("ncount") >= 50)), label
        assert st["boundary_list_overflows"] == int(np.sum(orc.field("bncount") >= 50)), label
        assert st["fluid_neighbours"] == int(orc.field("ncount").astype(np.int64).sum()), label
        assert st["boundary_neighbours"] == int(orc.field("bncount").astype(np.int64).sum()), label
        assert st["edge_reads"] == orc.edge_reads - before, label  # the oracle counts over its lifetime
    return sim, orc


def _clump(method):
    base = scenes.small_pcisph() if method == 1 else scenes.small_wcsph()
    sp = 0.02
    rng = Rng(31)
    vols = []
    for k in range(3):  # three lattices offset by a third of a spacing: ~3x the neighbours
        v = scenes._box(np.array([0.2, 0.16, 0.14], np.float32) * (1 - 1e-5),
                        np.array([0.1 + k * sp / 3, -0.2 + k * sp / 3, 0.02 + k * sp / 3], np.float32), sp / 2, rng)
        vols.append(v)
    return Scene("clump", dict(base.settings), base.container, vols, base.bodies)


@pytest.mark.parametrize("method", [1, 0])
def test_dense_clump_reaches_the_list_cap(method):
    sc = _clump(method)
    print("case clump", method, sc.n_fluid, flush=True)
    fields = (["density", "normal", "force_other", "pressure", "force_pressure", "pos", "vel"] if method == 1
              else ["density", "pressure", "normal", "accel", "pos", "vel", "vel_half"])
    sim, orc = _step_and_compare(sc, method, 2, fields, f"clump {method}")
    capped = int(np.sum(orc.field("ncount") >= 50))
    assert capped > sc.n_fluid // 4, capped


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 257, 1000])
def test_ragged_counts(n):
    base = scenes.small_pcisph()
    full = base.volumes[0]
    sc = Scene(f"ragged{n}", dict(base.settings), base.container, [full[:n].copy()], base.bodies)
    print("case ragged", n, flush=True)
    _step_and_compare(sc, 1, 2, ["density", "normal", "force_other", "pressure", "pos", "vel"], f"ragged {n}")


def test_no_fluid_particles():
    base = scenes.small_pcisph()
    sc = Scene("empty", dict(base.settings), base.container, [], base.bodies)
    sim = scenes.build(sc)
    assert sim.particle_count() == 0
    assert sim.boundary_particle_count() == sc.n_boundary
    sim.step(3)
    sim.synchronize()
    assert sim.positions().shape[0] == 0
    assert sim.step_stats()["nan_positions"] == 0


def test_relaxation_runs_every_correction_in_host_chunks():
    """The relaxation step's cap of 10000 corrections (pcisphsimluation.cpp:
    548-550) with an exit test that never passes (error ratio -1): the
    corrections past the third are enqueued eight at a time, the host
    reading the device's exit flag between chunks. All 10000 run, and the
    relaxed state and the next step (five corrections) equal the oracle's
    bit for bit."""
    base = scenes.small_pcisph(n_side=(8, 8, 6))
    settings = dict(base.settings)
    settings.update(pcisph_max_iterations=5, pcisph_error_ratio=-1.0)
    sc = Scene("relaxation_chunks", settings, base.container, base.volumes, base.bodies)
    print("case relaxation chunks", sc.n_fluid, flush=True)
    sim = scenes.build(sc)
    orc = build_oracle(sc)
    assert sim.step_stats()["iterations"] == 10000
    assert np.array_equal(sim.positions(), orc.field("pos"))
    sim.step(1)
    orc.step(1)
    assert sim.step_stats()["iterations"] == orc.iterations == 5
    assert np.array_equal(sim.positions(), orc.field("pos"))
    assert np.array_equal(sim.velocities(), orc.field("vel"))


def _bits_equal(a, b):
    """Bit equality of fp32 arrays, any NaN equal to any NaN."""
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | both_nan))


@pytest.mark.parametrize("method", [1, 0])
def test_nan_velocity_is_counted(method):
    """Failure detection forced from the host: one particle's velocity is set
    to NaN. Within the step every position is still finite (the hash, sort and
    lists see finite inputs); the NaN spreads through viscosity to the
    particle's neighbours and into their new positions, which the step
    statistics count (nan_positions) exactly as many as the oracle produces.
    Every output, NaN or not, equals the oracle's."""
    sc = scenes.small_pcisph() if method == 1 else scenes.small_wcsph()
    print("case nan velocity", method, sc.n_fluid, flush=True)
    sim = scenes.build(sc)
    orc = build_oracle(sc)
    pos, vel = sim.positions(), sim.velocities()
    vel[17, :3] = np.nan
    vh = sim.half_velocities() if method == 0 else None
    sim.set_velocities(vel)
    orc.set_state(pos, vel, vh)
    sim.step(1)
    orc.step(1)
    for f in INT_FIELDS:
        assert np.array_equal(sim.field(f), orc.field(f)), f
    for f in ["pos", "vel", "density", "pressure"]:
        assert _bits_equal(sim.field(f), orc.field(f)), f
    want = int(np.sum(~np.isfinite(orc.field("pos")[:, :3]).all(1)))
    assert want > 1
    assert sim.step_stats()["nan_positions"] == want
    # the state now holds NaN positions: this handle is not stepped again


def test_positions_outside_the_tank_grow_the_range_box():
    """Particles set far outside the container (but inside the 5 m periodic
    grid, one in the last cell column x = cps - 1, whose +x neighbour cells
    name other cells through the linear CELL_ID or lie past cps^3): the range
    box grows to hold them, nothing is out of the grid, and the step equals
    the oracle bit for bit, edge reads included."""
    sc = scenes.small_pcisph()
    print("case outside tank", sc.n_fluid, flush=True)
    sim = scenes.build(sc)
    orc = build_oracle(sc)
    pos, vel = sim.positions(), sim.velocities()
    pos[3, :3] = [2.2, 0.1, 0.0]
    pos[4, :3] = [-2.3, -0.3, 0.2]
    pos[5, :3] = [0.1, 0.2, 2.4]
    pos[6, :3] = [2.495, 2.49, 2.49]
    sim.set_positions(pos)
    orc.set_state(pos, vel)
    before = orc.edge_reads
    sim.step(1)
    orc.step(1)
    for f in INT_FIELDS + ["bkeys", "bperm"]:
        assert np.array_equal(sim.field(f), orc.field(f)), f
    for f in ["ranges", "branges"]:
        assert np.array_equal(sim.field(f), orc.field(f)), f
    for f in ["density", "pressure", "force_pressure", "pos", "vel"]:
        assert _bits_equal(sim.field(f), orc.field(f)), f
    st = sim.step_stats()
    assert st["out_of_grid"] == 0
    assert st["edge_reads"] == orc.edge_reads - before > 0
