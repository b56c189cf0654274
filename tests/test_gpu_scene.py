"""Scene files through sph_scene_* (Scene::load_scene, scene/scene.cpp:322-476).

* many_objects.json (two cubes and a sphere in a double dam break) with
  config.pci: the scene holds the reference's bodies with the samplers'
  counts, the fluid holds the box volumes' lattices, and its frames are
  bit-identical to the same scene assembled by hand through the Python API
  (set_rect_limits, add_volume, add_boundary, a rigid world, coupled_step).
* A model body (the teapot of the reference's data) loads, samples as
  sph_sample_model does, and steps without non-finite positions.
"""
import json
import os

import numpy as np
import pytest

from fluids_amd import (FluidSimulationFactory, box_volume_particles, cube_sampling, load_scene, load_settings,
                        model_sampling, sphere_sampling)
from fluids_amd.rigid import RigidWorld, coupled_step
from fluids_amd.scenes import quat_to_rows

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_many_objects_scene_matches_manual_assembly():
    phys, sim = load_settings(os.path.join(GOLDEN, "config.pci"))
    sc = load_scene(os.path.join(GOLDEN, "many_objects.json"), phys, sim)
    r = sim.fluid_particle_radius
    bodies = sc.bodies()
    assert [b["type"] for b in bodies] == ["cube", "cube", "sphere"]
    assert [b["name"] for b in bodies] == ["cube_1", "cube_2", "sphere_1"]
    assert bodies[0]["samples"] == cube_sampling(0.3, 2 * r).shape[0]
    assert bodies[1]["samples"] == cube_sampling(0.2, 2 * r).shape[0]
    assert bodies[2]["samples"] == sphere_sampling(0.1, 2 * r).shape[0]
    v1 = box_volume_particles([0.5, 1.0, 0.7], [0.74, 0.0, -0.1], r)
    v2 = box_volume_particles([0.5, 1.0, 0.7], [-0.74, 0.0, 0.1], r)
    assert sc.simulation.particle_count() == v1.shape[0] + v2.shape[0]

    # the same scene by hand
    fs = FluidSimulationFactory.build_simulation(phys, sim)
    fs.set_rect_limits(2.0, 1.0, 1.0)
    fs.add_volume(v1)
    fs.add_volume(v2)
    w = RigidWorld(-9.8)
    w.set_container(2.0, 1.0, 1.0)
    pairs = []
    specs = [(cube_sampling(0.3, 2 * r), 12.0, [0.0, -0.35, 0.0], [0, 1, 0, 1], ("box", 0.15)),
             (cube_sampling(0.2, 2 * r), 3.0, [0.0, -0.1, 0.0], [0, 1, 0, 1], ("box", 0.1)),
             (sphere_sampling(0.1, 2 * r), 0.2, [0.0, 0.1, 0.0], [0, 0, 0, 1], ("sphere", 0.1))]
    for samples, mass, com, q, (kind, size) in specs:
        sid = fs.add_boundary(samples, True, mass, quat_to_rows(q), np.array(com, np.float32))
        rid = w.add_box(mass, (size, size, size), com, q) if kind == "box" else w.add_sphere(mass, size, com)
        pairs.append((sid, rid))
    assert np.array_equal(fs.positions(), sc.simulation.positions())
    for frame in range(12):
        sc.step(1)
        coupled_step(fs, w, pairs)
        assert np.array_equal(fs.positions(), sc.simulation.positions()), frame
    for k in range(3):
        assert np.array_equal(w.state(pairs[k][1])["com"], sc.world.state(bodies[k]["rigid_id"])["com"])
    assert sc.simulation.step_stats()["nan_positions"] == 0


def test_model_body_scene(tmp_path):
    phys, sim = load_settings(os.path.join(GOLDEN, "config.pci"))
    scene = {
        "fluid_volumes": [{"type": "box", "name": "f", "size": [0.4, 0.5, 0.5], "center": [0.7, -0.2, 0.0]}],
        "container": {"width": 2.0, "height": 1.0, "depth": 1.0},
        "rigid_bodies": [{"type": "model", "name": "teapot", "obj_filename": "teapot.obj", "mass": 2.0,
                          "center": [-0.2, -0.3, 0.0], "rotation": [0, 0, 0, 1]}],
    }
    path = tmp_path / "teapot_scene.json"
    path.write_text(json.dumps(scene))
    sc = load_scene(str(path), phys, sim, base_dir=GOLDEN)
    b = sc.bodies()
    assert b[0]["type"] == "model"
    assert b[0]["samples"] == model_sampling(os.path.join(GOLDEN, "teapot.obj"), 2 * sim.fluid_particle_radius).shape[0]
    sc.step(10)
    p = sc.simulation.positions()
    assert np.isfinite(p).all()
    assert sc.simulation.step_stats()["nan_positions"] == 0
