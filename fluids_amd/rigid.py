"""Python side of the host rigid-body world (include/sph.h, sph_rigid_*),
which takes Bullet's place in the two-way coupling loop of
scene/scene.cpp:100-110: the fluid steps, each body with mass > 0 receives the
fluid's force and torque (rigidbody.cpp:49-54), then the world steps by dt.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib

_F3 = C.c_float * 3
_F4 = C.c_float * 4
_F9 = C.c_float * 9


def _v(a, n):
    if a is None:
        return None
    a = np.asarray(a, dtype=np.float32).reshape(-1)
    if a.size != n:
        raise ValueError(f"expected {n} numbers, got {a.size}")
    return (C.c_float * n)(*a.tolist())


class RigidWorld:
    """btDiscreteDynamicsWorld with gravity (0, gravity_y, 0) (scene.cpp:304-316)."""

    def __init__(self, gravity_y: float = -9.8):
        L = lib()
        self._h = L.sph_rigid_create(float(gravity_y))
        if not self._h:
            raise RuntimeError(L.sph_last_error().decode())

    @classmethod
    def _borrow(cls, handle):
        """A view of a world owned elsewhere (a loaded scene); close() leaves it."""
        obj = cls.__new__(cls)
        obj._h = handle
        obj._borrowed = True
        return obj

    def close(self):
        if getattr(self, "_h", None):
            if not getattr(self, "_borrowed", False):
                lib().sph_rigid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_container(self, width, height, depth, center=None):
        """FishTank collision mesh: open-top box (fishtank.cpp:143-182)."""
        check(lib().sph_rigid_set_container(self._h, width, height, depth, _v(center, 3)))

    def add_box(self, mass, half, pos, quat=(0, 0, 0, 1)) -> int:
        return check(lib().sph_rigid_add_box(self._h, mass, _v(half, 3), _v(pos, 3), _v(quat, 4)))

    def add_cube(self, side, mass, pos, quat=(0, 0, 0, 1)) -> int:
        """Cube(side, mass, center, rotation) (scene/cube.cpp:129-134)."""
        h = side / 2.0
        return self.add_box(mass, (h, h, h), pos, quat)

    def add_sphere(self, mass, radius, pos) -> int:
        return check(lib().sph_rigid_add_sphere(self._h, mass, radius, _v(pos, 3)))

    def add_static(self, pos, quat=(0, 0, 0, 1)) -> int:
        return check(lib().sph_rigid_add_static(self._h, _v(pos, 3), _v(quat, 4)))

    def apply_force_n_torque(self, body, f, t):
        check(lib().sph_rigid_apply_force_n_torque(self._h, body, _v(f, 3), _v(t, 3)))

    def step(self, dt):
        check(lib().sph_rigid_step(self._h, dt))

    def state(self, body) -> dict:
        R, com, q, lin, ang = _F9(), _F3(), _F4(), _F3(), _F3()
        check(lib().sph_rigid_get_state(self._h, body, R, com, q, lin, ang))
        return {"R": np.array(R[:], np.float32).reshape(3, 3), "com": np.array(com[:], np.float32),
                "quat": np.array(q[:], np.float32), "lin": np.array(lin[:], np.float32),
                "ang": np.array(ang[:], np.float32)}

    def set_state(self, body, pos=None, quat=None, lin=None, ang=None):
        check(lib().sph_rigid_set_state(self._h, body, _v(pos, 3), _v(quat, 4), _v(lin, 3), _v(ang, 3)))

    def body_count(self) -> int:
        return check(lib().sph_rigid_body_count(self._h))


def coupled_step(sim, world: RigidWorld, pairs):
    """One frame of Scene::render: pairs is a list of (fluid body id, rigid
    body id); body states go to the fluid, the fluid steps, wrenches go to
    bodies with mass > 0, then the world steps by the fluid's time step."""
    n = len(pairs)
    a = (C.c_int * max(n, 1))(*[p[0] for p in pairs])
    b = (C.c_int * max(n, 1))(*[p[1] for p in pairs])
    check(lib().sph_coupled_step(sim._h, world.handle, n, a, b))
