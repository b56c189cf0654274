"""CPU tests of the mesh-body samplers and the scene-file reader's inputs.

* Sphere::surface_sampling: the library's samples equal, bit for bit, the
  oracle's numpy float32 restatement (oracle/mesh_samplers.py), whose
  generator is pinned to the C++ standard's known answer for minstd_rand0;
  the icosphere's vertices lie on the sphere and number 12 + 3 * (20 + 80 +
  320) = 1272.
* Model::surface_sampling: the samples equal, in count, order and bits, a
  float32 restatement of the reference's voxelizer (oracle/mesh_samplers.py
  voxelize, with its own separating-axis test) on the reference's teapot and
  bunny models (committed fixtures); that overlap test is in turn pinned to
  float64 clipping of each triangle against each candidate box.
"""
import os

import numpy as np
import pytest

from fluids_amd import model_sampling, sphere_sampling
from oracle import mesh_samplers as ms

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_minstd_rand0_known_answer():
    g = ms.MinStdRand0()
    assert [g() for _ in range(3)] == [16807, 282475249, 1622650073]
    g = ms.MinStdRand0()
    for _ in range(9999):
        g()
    assert g() == 1043618065


def test_icosphere_invariants():
    V, I = ms.icosphere(0.25, 3)
    v = V.reshape(-1, 3)
    assert v.shape[0] == 12 + 3 * (20 + 80 + 320) and I.size == 3 * 1280
    r = np.linalg.norm(v.astype(np.float64), axis=1)
    assert np.abs(r - 0.25).max() < 1e-6


def test_sphere_sampling_matches_restatement():
    for radius, spacing in [(0.1, 0.02), (0.05, 0.01)]:
        a = sphere_sampling(radius, spacing)
        b = ms.sphere_sampling(radius, spacing)
        assert a.shape == b.shape, (a.shape, b.shape)
        assert np.array_equal(a, b)
        # the mesh vertices close the list and lie on the sphere
        tail = np.linalg.norm(a[-1272:, :3].astype(np.float64), axis=1)
        assert np.abs(tail - radius).max() < 1e-6 * max(radius, 1)


@pytest.mark.parametrize("model,spacing", [("teapot.obj", 0.02), ("teapot.obj", 0.05), ("bunny.obj", 0.01),
                                           ("bunny.obj", 0.02)])
def test_model_sampling_exact(model, spacing):
    """Model::surface_sampling (scene/model.cpp:170-214 over vx_voxelize,
    voxelizer.h:629-760) on two of the reference's models: the library's
    samples equal the float32 restatement's (oracle/mesh_samplers.py
    voxelize) in count, order and bits; 8 corners per voxel, each the
    voxel centre +- s/2."""
    path = os.path.join(GOLDEN, model)
    a = model_sampling(path, spacing)
    verts, tris = ms.read_obj(path)
    b = ms.voxelize(verts, tris, spacing, spacing / 10)
    print(f"{model} at {spacing}: {a.shape[0]} samples ({a.shape[0] // 8} voxels)")
    assert a.shape == b.shape and a.shape[0] > 0 and a.shape[0] % 8 == 0
    assert np.array_equal(a, b)
    corners = a[:, :3].reshape(-1, 8, 3).astype(np.float64)
    centres = corners.mean(axis=1)
    assert np.abs(np.abs(corners - centres[:, None, :]) - spacing / 2).max() < 1e-6


def test_voxel_overlap_test_against_clipping():
    """The restatement's separating-axis test (the one the samples above are
    pinned to) against float64 clipping of each triangle to each candidate
    box of the teapot at 0.02 m: every box that the triangle meets when
    shrunk by 1e-6 is kept, every box it misses when grown by 1e-6 is not,
    with no disagreement (boxes within 1e-6 of touching are counted apart)."""
    verts, tris = ms.read_obj(os.path.join(GOLDEN, "teapot.obj"))
    agree_in, agree_out, disagree, ambiguous = ms.sat_against_clipping(verts, tris, 0.02, 0.002)
    print(f"overlap {agree_in}, separated {agree_out}, disagreeing {disagree}, within rounding {ambiguous}")
    assert disagree == 0
    assert agree_in > 5000 and agree_out > 5000
    assert ambiguous <= (agree_in + agree_out) // 1000
