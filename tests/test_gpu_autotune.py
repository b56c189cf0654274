"""Launch shapes of the single-scene PCISPH kernels (sph_autotune,
sph_set_launch): the reference autotunes its work-group sizes
(opencl/clkernel.cpp:85-151); here threads per block and, for the correction
kernels, grid-stride grids. A shape must never change a result: every shape
gives the same bits as the default and as the oracle, and tuning restores
the state it started from."""
import numpy as np
import pytest

from fluids_amd import scenes
from helpers import build_oracle

pytestmark = pytest.mark.gpu

KERNELS = ("normals", "forces_predict", "update_pressure", "pressure_force")
FIELDS = ["density", "normal", "force_other", "pressure", "err", "force_pressure", "pos_pred", "pos", "vel"]


def _run(sim, pos, vel, steps=2):
    sim.set_positions(pos)
    sim.set_velocities(vel)
    sim.step(steps)
    return {f: sim.field(f) for f in FIELDS}


def test_launch_shapes_do_not_change_results():
    sc = scenes.small_pcisph(n_side=(40, 25, 50), spacing=0.005, cube=True)
    print("case launch shapes", sc.n_fluid, flush=True)
    sim = scenes.build(sc)
    pos, vel = sim.positions(), sim.velocities()
    ref = _run(sim, pos, vel)
    orc = build_oracle(sc)
    orc.set_state(pos, vel)
    orc.step(2)
    for f in FIELDS:
        assert np.array_equal(ref[f], orc.field(f)), f
    shapes = [(64, 0), (128, 0), (256, 0)]
    for k in KERNELS:
        cands = shapes + ([(64, 256), (128, 1024), (256, 4096)] if k in ("update_pressure", "pressure_force") else [])
        for b, g in cands:
            sim.set_launch(k, b, g)
            got = _run(sim, pos, vel)
            for f in FIELDS:
                assert np.array_equal(got[f], ref[f]), (k, b, g, f)
        sim.set_launch(k, 256, 0)


def test_autotune_restores_state_and_caches(tmp_path):
    sc = scenes.small_pcisph(n_side=(40, 25, 50), spacing=0.005)
    sim = scenes.build(sc)
    sim.step(2)
    pos, vel = sim.positions(), sim.velocities()
    cache = str(tmp_path / "tune.txt")
    assert sim.autotune(cache, trials=1) is False
    assert np.array_equal(sim.positions(), pos) and np.array_equal(sim.velocities(), vel)
    chosen = {k: sim.launch(k) for k in KERNELS}
    print("tuned shapes", chosen, flush=True)
    sim2 = scenes.build(sc)
    assert sim2.autotune(cache, trials=1) is True
    assert {k: sim2.launch(k) for k in KERNELS} == chosen
    # tuned and default shapes step to the same bits
    sim2.set_positions(pos)
    sim2.set_velocities(vel)
    for k in KERNELS:
        sim.set_launch(k, 256, 0)
    sim.set_positions(pos)
    sim.set_velocities(vel)
    sim.step(3)
    sim2.step(3)
    assert np.array_equal(sim.positions(), sim2.positions())
