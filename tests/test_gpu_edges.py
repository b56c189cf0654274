"""Edge cases of the HIP path against the CPU oracle, bit for bit:

* a dense clump (three interleaved lattices) where most lists reach the
  cap of 50 (NEIGH_LIST_MAX_LENGTH, kernels/common.h:4; the walk stops at the
  cap, grid.cl:170-177), for PCISPH and WCSPH, with the overflow and
  edge-read counters;
* ragged particle counts around the wave (64) and workgroup (256) sizes,
  down to a single particle;
* a step with no fluid particles at all (only the tank).
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
        assert st["edge_reads"] == orc.edge_reads, label
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


def test_correction_barrier_timeout_is_reported():
    """Failure handling of the grid barriers in the PCISPH correction loop,
    forced from the host: with a zero wait bound every workgroup but the
    last to arrive gives up at its first barrier, the launch ends, and the
    step statistics raise instead of returning a silently wrong step. The
    exit test is made to fail (error ratio -1) so that corrections past the
    third run; 960 particles give the launch four workgroups."""
    from fluids_amd._lib import SphError
    base = scenes.small_pcisph()
    settings = dict(base.settings)
    settings.update(pcisph_max_iterations=5, pcisph_error_ratio=-1.0)
    sc = Scene("barrier_timeout", settings, base.container, base.volumes, base.bodies)
    print("case barrier timeout", sc.n_fluid, flush=True)
    sim = scenes.build(sc)
    assert sim.step_stats()["iterations"] == 10000  # the relaxation step ran every correction
    sim.step(1)
    assert sim.step_stats()["iterations"] == 5
    sim.set_barrier_timeout(0.0)
    sim.step(1)
    sim.synchronize()
    with pytest.raises(SphError, match="grid barrier"):
        sim.step_stats()
