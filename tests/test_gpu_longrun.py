"""Long runs of the HIP path against the CPU oracle, through the C ABI.

* The 32 000-particle WCSPH dam break over its full 1000 steps: positions are
  compared bit for bit every 100 steps, and the kinetic energy and the mean
  density error are checked against the oracle's within 0.5 %.
* The PCISPH double dam break with the 12 kg cube, two-way coupled to the host
  rigid-body world (sph_coupled_step), at a 0.02 m lattice (87 500 particles)
  over 300 steps and at its configured 0.01 m lattice (700 000 particles)
  over 500 steps: the oracle runs the same loop with its own world. Both
  sides add the body wrench in the same pinned order, so the fluid positions
  stay bit-equal and the two worlds receive the same wrenches; the cube's
  trajectory and the fluid's kinetic energy and mean density error are also
  checked against the 0.5 % contract.
"""
import os

import numpy as np
import pytest

import oracle
from fluids_amd import scenes
from fluids_amd.rigid import RigidWorld, coupled_step
from helpers import build_oracle

pytestmark = pytest.mark.gpu


def _threads():
    oracle.set_threads(min(16, os.cpu_count() or 1))


def _stats(vel, density, rho0):
    ke = float(np.sum(np.asarray(vel[:, :3], np.float64) ** 2))
    derr = float(np.mean(np.abs(np.asarray(density, np.float64) - rho0)) / rho0)
    return ke, derr


def test_wcsph_dambreak_1000_steps():
    _threads()
    sc = scenes.c1_wcsph_dambreak()
    sim = scenes.build(sc)
    orc = build_oracle(sc)
    rho0 = sc.settings["rest_density"]
    for block in range(10):
        sim.step(100)
        orc.step(100)
        label = f"step {100 * (block + 1)}"
        pg, po = sim.positions(), orc.field("pos")
        assert np.array_equal(pg, po), f"{label}: positions differ in {np.sum(pg != po)} values"
        ke_g, de_g = _stats(sim.velocities(), sim.field("density"), rho0)
        ke_o, de_o = _stats(orc.field("vel"), orc.field("density"), rho0)
        print(f"{label}: KE {ke_g:.6e} / {ke_o:.6e}, mean density error {de_g:.4e} / {de_o:.4e}", flush=True)
        assert abs(ke_g - ke_o) <= 5e-3 * max(ke_o, 1e-12), label
        assert abs(de_g - de_o) <= 5e-3 * max(de_o, 1e-12), label
    st = sim.step_stats()
    assert st["nan_positions"] == 0 and st["out_of_grid"] == 0


def _world(sc):
    w = RigidWorld(-9.8)
    w.set_container(*sc.container)
    cube = sc.bodies[0]
    rid = w.add_cube(0.3, cube.mass, cube.com, (0, 1, 0, 1))
    return w, rid


@pytest.mark.parametrize("spacing,steps,every", [(0.02, 300, 50), (0.01, 500, 100)])
def test_pcisph_double_dambreak_coupled_cube(spacing, steps, every):
    """0.02 m (87 500 particles, 300 steps) and the configured run: 0.01 m
    lattice (700 000 particles), the 12 kg cube sampled at 0.01 m, 500 steps
    (data/scenes/double_dambreak.json, data/scenes/many_objects.json:31-37)."""
    _threads()
    sc = scenes.c2_pcisph_double_dambreak(spacing=spacing)
    print("case double dam break", spacing, sc.n_fluid, sc.n_boundary, steps, flush=True)
    sim = scenes.build(sc)
    orc = build_oracle(sc)
    wg, rg = _world(sc)
    wo, ro = _world(sc)
    rho0 = sc.settings["rest_density"]
    dt = sc.settings["time_step"]
    start = wg.state(rg)["com"].astype(np.float64)
    max_dev = 0.0
    for step in range(1, steps + 1):
        coupled_step(sim, wg, [(0, rg)])
        # the oracle runs Scene::render's order in Python with its own world
        s = wo.state(ro)
        orc.set_body_state(0, s["R"], s["com"], s["lin"], s["ang"])
        orc.step(1)
        f, t, _, _ = orc.body_wrench(0)
        wo.apply_force_n_torque(ro, f, t)
        wo.step(dt)
        if step % every == 0:
            pg, po = sim.positions(), orc.field("pos")
            assert np.array_equal(pg, po), f"step {step}: positions differ in {np.sum(pg != po)} values"
            cg, co = wg.state(rg), wo.state(ro)
            moved = max(np.linalg.norm(co["com"] - start), 1e-3)
            dev = float(np.linalg.norm(cg["com"].astype(np.float64) - co["com"]) / moved)
            qdev = float(np.abs(cg["quat"] - co["quat"]).max())
            max_dev = max(max_dev, dev)
            ke_g, de_g = _stats(sim.velocities(), sim.field("density"), rho0)
            ke_o, de_o = _stats(orc.field("vel"), orc.field("density"), rho0)
            print(f"step {step}: cube {cg['com']} vs {co['com']} (moved {moved:.4f} m, dev {dev:.2e}, "
                  f"quat dev {qdev:.2e}); KE {ke_g:.6e}/{ke_o:.6e}; density err {de_g:.4e}/{de_o:.4e}", flush=True)
            assert dev <= 5e-3, step
            assert qdev <= 5e-3, step
            assert abs(ke_g - ke_o) <= 5e-3 * max(ke_o, 1e-12), step
            assert abs(de_g - de_o) <= 5e-3 * max(de_o, 1e-12), step
    # the fluid reached the cube: it was pushed off its start
    assert np.linalg.norm(wg.state(rg)["com"] - start) > 1e-3
