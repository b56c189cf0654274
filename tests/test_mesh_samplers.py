"""CPU tests of the mesh-body samplers and the scene-file reader's inputs.

* Sphere::surface_sampling: the library's samples equal, bit for bit, the
  oracle's numpy float32 restatement (oracle/mesh_samplers.py), whose
  generator is pinned to the C++ standard's known answer for minstd_rand0;
  the icosphere's vertices lie on the sphere and number 12 + 3 * (20 + 80 +
  320) = 1272.
* Model::surface_sampling: 8 corners per voxel; the voxel set equals the one
  found by clipping each triangle against each candidate box in float64
  (a different overlap algorithm from the library's separating axes),
  on the teapot model of the reference's data (a committed fixture).
"""
import os

import numpy as np

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


def test_model_sampling_voxel_set():
    path = os.path.join(GOLDEN, "teapot.obj")
    s = 0.02
    a = model_sampling(path, s)
    assert a.shape[0] % 8 == 0 and a.shape[0] > 0
    corners = a[:, :3].reshape(-1, 8, 3).astype(np.float64)
    centres = corners.mean(axis=1)
    # each voxel's 8 corners are the centre +- s/2
    assert np.abs(np.abs(corners - centres[:, None, :]) - s / 2).max() < 1e-6
    # voxel centres come from float loops started at each triangle's snapped
    # corner, so one lattice voxel can be met at centres a few ulp apart and
    # (beyond the 1e-7 tolerance of the voxelizer's table) kept twice, as in
    # the reference
    idx = np.rint(centres / s).astype(int)
    assert np.abs(centres - idx * s).max() < 1e-5
    mine = {tuple(i) for i in idx}
    verts, tris = ms.read_obj(path)
    ref = ms.voxel_centres(verts, tris, s, s / 10)
    diff = mine ^ ref
    print(f"voxels: library {len(mine)} ({centres.shape[0]} with repeats), clipping {len(ref)}, "
          f"differing {len(diff)}")
    assert len(diff) <= max(2, len(ref) // 500)
