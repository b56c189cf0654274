"""Python side of batched WCSPH (include/sph.h, sph_batch_*): many
independent scenes packed into one particle array and advanced together, one
HIP graph per step."""
import ctypes as C

import numpy as np

from ._lib import SphSettings, check, lib
from .simulation import settings_from_dict, make_settings

KERNELS = ["batch_hash", "batch_gather_ranges", "batch_lists_density", "batch_normals", "batch_accel_integrate"]


def _settings(s) -> SphSettings:
    if isinstance(s, SphSettings):
        return s
    phys, sim = settings_from_dict(s)
    return make_settings(phys, sim)


class BatchSimulation:
    def __init__(self):
        L = lib()
        self._h = L.sph_batch_create()
        if not self._h:
            raise RuntimeError(L.sph_last_error().decode())
        self._keep = []

    def close(self):
        if getattr(self, "_h", None):
            lib().sph_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, settings, container, particles) -> int:
        p = np.ascontiguousarray(particles, np.float32).reshape(-1, 4)
        st = _settings(settings)
        return check(lib().sph_batch_add_scene(self._h, C.byref(st), float(container[0]), float(container[1]),
                                               float(container[2]), p.ctypes.data, p.shape[0]))

    def finalize(self):
        check(lib().sph_batch_finalize(self._h))

    def autotune(self, cache=None, trials=4) -> bool:
        rc = check(lib().sph_batch_autotune(self._h, cache.encode() if cache else None, int(trials)))
        return rc == 1

    def launch(self, kernel) -> tuple:
        k = KERNELS.index(kernel) if isinstance(kernel, str) else int(kernel)
        b, g, s = C.c_int(), C.c_int(), C.c_int()
        check(lib().sph_batch_get_launch(self._h, k, C.byref(b), C.byref(g), C.byref(s)))
        return b.value, g.value, s.value

    def set_launch(self, kernel, block, grid, stage):
        k = KERNELS.index(kernel) if isinstance(kernel, str) else int(kernel)
        check(lib().sph_batch_set_launch(self._h, k, block, grid, stage))

    def set_graph(self, on: bool):
        check(lib().sph_batch_set_graph(self._h, int(bool(on))))

    def step(self, n: int = 1):
        check(lib().sph_batch_step(self._h, int(n)))

    def synchronize(self):
        check(lib().sph_batch_synchronize(self._h))

    @property
    def scene_count(self) -> int:
        return check(lib().sph_batch_scene_count(self._h))

    @property
    def total_particles(self) -> int:
        return check(lib().sph_batch_total_particles(self._h))

    def particle_count(self, scene) -> int:
        return check(lib().sph_batch_particle_count(self._h, scene))

    def _get4(self, fn, scene):
        n = self.particle_count(scene)
        out = np.zeros((n, 4), np.float32)
        check(fn(self._h, scene, out.ctypes.data, n))
        return out

    def positions(self, scene):
        return self._get4(lib().sph_batch_get_positions, scene)

    def velocities(self, scene):
        return self._get4(lib().sph_batch_get_velocities, scene)

    def half_velocities(self, scene):
        return self._get4(lib().sph_batch_get_half_velocities, scene)

    def set_positions(self, scene, a):
        a = np.ascontiguousarray(a, np.float32).reshape(-1, 4)
        check(lib().sph_batch_set_positions(self._h, scene, a.ctypes.data, a.shape[0]))

    def set_velocities(self, scene, a):
        a = np.ascontiguousarray(a, np.float32).reshape(-1, 4)
        check(lib().sph_batch_set_velocities(self._h, scene, a.ctypes.data, a.shape[0]))

    def density(self, scene):
        n = self.particle_count(scene)
        out = np.zeros(n, np.float32)
        check(lib().sph_batch_get_density(self._h, scene, out.ctypes.data, n))
        return out

    def neighbour_counts(self, scene):
        n = self.particle_count(scene)
        out = np.zeros(n, np.int32)
        check(lib().sph_batch_get_neighbour_counts(self._h, scene, out.ctypes.data, n))
        return out

    def set_profiling(self, on=True, kernel=None):
        check(lib().sph_batch_set_profiling(self._h, int(bool(on)), kernel.encode() if kernel else None))

    def kernel_times(self) -> dict:
        cap = 64
        names = C.create_string_buffer(48 * cap)
        launches = (C.c_int * cap)()
        ms = (C.c_double * cap)()
        n = check(lib().sph_batch_get_kernel_times(self._h, names, launches, ms, cap))
        out = {}
        for k in range(min(n, cap)):
            name = names.raw[48 * k:48 * (k + 1)].split(b"\0")[0].decode()
            out[name] = (launches[k], ms[k])
        return out
