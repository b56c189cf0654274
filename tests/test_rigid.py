"""CPU tests of the host rigid-body world that replaces Bullet in the coupling
loop (scene/scene.cpp:100-110, 304-316; rigidbody.cpp:49-54). Bullet is not
available, so each behaviour is pinned by a closed form or an invariant:
semi-implicit Euler free fall, the impulse of a torque on a cube's known
inertia, apply_force_n_torque replacing (not adding to) earlier forces and the
step clearing them, resting contact on the tank floor, walls that hold a
thrown sphere, and a stack of two cubes."""
import numpy as np
import pytest

from fluids_amd import SphError
from fluids_amd.rigid import RigidWorld
from fluids_amd.scenes import quat_to_rows

G = -9.8


def test_free_fall_closed_form():
    # v_n = g n dt, y_n = y0 + g dt^2 n (n + 1) / 2 (velocity first, then position)
    w = RigidWorld(G)
    b = w.add_sphere(2.0, 0.1, (0.0, 5.0, 0.0))
    dt, n = 1e-3, 400
    for _ in range(n):
        w.step(dt)
    s = w.state(b)
    assert s["lin"][1] == pytest.approx(G * n * dt, rel=1e-4)
    assert s["com"][1] == pytest.approx(5.0 + G * dt * dt * n * (n + 1) / 2, rel=1e-4)
    assert s["com"][0] == 0.0 and s["com"][2] == 0.0
    assert np.allclose(s["quat"], [0, 0, 0, 1])


def test_force_replaces_and_step_clears():
    w = RigidWorld(G)
    b = w.add_sphere(3.0, 0.1, (0.0, 0.0, 0.0))
    # first a wrong force, then the one that balances gravity: the second call
    # replaces the first (clearForces, applyCentralForce)
    w.apply_force_n_torque(b, (100.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    w.apply_force_n_torque(b, (0.0, 3.0 * 9.8, 0.0), (0.0, 0.0, 0.0))
    w.step(1e-3)
    s = w.state(b)
    assert abs(s["lin"][1]) < 1e-6 and s["lin"][0] == 0.0
    # the step cleared it: the next step falls freely
    w.step(1e-3)
    assert w.state(b)["lin"][1] == pytest.approx(G * 1e-3, rel=1e-5)


def test_torque_on_cube_inertia():
    # cube of side a, mass m: I = m a^2 / 6 about every axis
    m, a, tau, dt = 12.0, 0.3, 0.9, 1e-3
    w = RigidWorld(0.0)
    b = w.add_cube(a, m, (0, 0, 0))
    w.apply_force_n_torque(b, (0, 0, 0), (0, tau, 0))
    w.step(dt)
    om = tau * dt / (m * a * a / 6)
    s = w.state(b)
    assert s["ang"][1] == pytest.approx(om, rel=1e-5)
    n = 500
    for _ in range(n):
        w.step(dt)
    q = w.state(b)["quat"]
    angle = 2 * np.arctan2(np.linalg.norm(q[:3]), q[3])
    assert angle == pytest.approx(om * dt * (n + 1), rel=1e-4)
    assert abs(q[0]) < 1e-6 and abs(q[2]) < 1e-6


def test_rotation_rows_match_scene_quaternion():
    w = RigidWorld(G)
    q = (0.0, 1.0, 0.0, 1.0)  # many_objects.json: unnormalised, as Bullet takes it
    b = w.add_cube(0.3, 12.0, (0.0, -0.35, 0.0), q)
    R = w.state(b)["R"].reshape(9)
    assert np.allclose(R, quat_to_rows(q), atol=1e-6)
    assert np.allclose(R.reshape(3, 3) @ R.reshape(3, 3).T, np.eye(3), atol=1e-6)


def test_cube_comes_to_rest_on_floor():
    w = RigidWorld(G)
    w.set_container(2.0, 1.0, 1.0)
    b = w.add_cube(0.3, 12.0, (0.1, -0.2, 0.05), (0, 1, 0, 1))
    for _ in range(3000):
        w.step(1e-3)
    s = w.state(b)
    bottom = s["com"][1] - 0.15
    assert abs(bottom - (-0.5)) < 5e-3, bottom
    assert np.linalg.norm(s["lin"]) < 1e-2 and np.linalg.norm(s["ang"]) < 1e-1
    assert abs(s["com"][0] - 0.1) < 1e-3 and abs(s["com"][2] - 0.05) < 1e-3


def test_walls_hold_a_thrown_sphere():
    w = RigidWorld(G)
    w.set_container(2.0, 1.0, 1.0)
    b = w.add_sphere(0.2, 0.05, (0.0, 0.0, 0.0))
    w.set_state(b, lin=(6.0, 0.0, -4.0))
    for _ in range(2000):
        w.step(1e-3)
        c = w.state(b)["com"]
        assert -1.0 + 0.05 - 0.01 <= c[0] <= 1.0 - 0.05 + 0.01
        assert -0.5 + 0.05 - 0.01 <= c[2] <= 0.5 - 0.05 + 0.01
        assert c[1] >= -0.5 + 0.05 - 0.01


def test_two_cubes_stack():
    w = RigidWorld(G)
    w.set_container(2.0, 1.0, 1.0)
    lo = w.add_cube(0.3, 12.0, (0.0, -0.34, 0.0))
    hi = w.add_cube(0.2, 3.0, (0.0, -0.08, 0.0))
    for _ in range(3000):
        w.step(1e-3)
    a, b = w.state(lo)["com"], w.state(hi)["com"]
    assert abs((a[1] - 0.15) - (-0.5)) < 5e-3
    assert abs((b[1] - 0.10) - (a[1] + 0.15)) < 5e-3


def test_static_bodies_and_errors():
    w = RigidWorld(G)
    s = w.add_static((0.0, 0.2, 0.0), (0, 0, 0, 1))
    w.step(1e-3)
    assert np.allclose(w.state(s)["com"], [0.0, 0.2, 0.0])
    with pytest.raises(SphError):
        w.step(0.0)
    with pytest.raises(SphError):
        w.state(7)
    with pytest.raises(SphError):
        w.add_cube(0.3, -1.0, (0, 0, 0))
    with pytest.raises(SphError):
        w.set_container(2.0, 0.0, 1.0)
    assert w.body_count() == 1
