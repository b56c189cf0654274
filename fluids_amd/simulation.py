"""Python mirror of the reference's simulation interface over libsph.

Names follow fluid/simulation/fluidsimulation.h:23-92 (FluidSimulation and
its PCISPH/WCSPH implementations), fluidsimulationfactory.h:13-40 and
settings/{physics,simulation}settings.h, so code written against the
reference's solver reads the same. Every call goes through the C ABI of
include/sph.h; errors raise SphError.
"""
import ctypes as C
from dataclasses import asdict, dataclass, field

import numpy as np

from ._lib import (SphError, FIELDS, I32_FIELDS, LIST_CAP, PCISPH, U32_FIELDS, VEC4_FIELDS, WCSPH,
                   SphSettings, SphStepStats, check, lib)


@dataclass
class PhysicsSettings:
    """settings/physicssettings.h:5-25 (defaults of the reference)."""
    rest_density: float = 998.29
    k_viscosity: float = 0.0035105
    gravity: float = 9.8
    gas_stiffness: float = 3.0
    surface_tension: float = 0.0728


@dataclass
class SimulationSettings:
    """settings/simulationsettings.h:9-39 (defaults of the reference)."""
    time_step: float = 0.01
    max_vel: float = 50.0
    fluid_particle_radius: float = 0.008
    fluid_support_radius: float = 0.032
    pcisph_max_iterations: int = 7
    pcisph_error_ratio: float = 0.01
    sim_method: int = WCSPH
    flags: int = 0


def make_settings(physics: PhysicsSettings, sim: SimulationSettings) -> SphSettings:
    s = SphSettings()
    s.time_step = sim.time_step
    s.max_vel = sim.max_vel
    s.fluid_particle_radius = sim.fluid_particle_radius
    s.fluid_support_radius = sim.fluid_support_radius
    s.pcisph_max_iterations = sim.pcisph_max_iterations
    s.pcisph_error_ratio = sim.pcisph_error_ratio
    s.method = sim.sim_method
    s.rest_density = physics.rest_density
    s.k_viscosity = physics.k_viscosity
    s.gravity = physics.gravity
    s.gas_stiffness = physics.gas_stiffness
    s.surface_tension = physics.surface_tension
    s.flags = sim.flags
    return s


def settings_from_dict(d: dict):
    """(PhysicsSettings, SimulationSettings) from a flat dict of config keys."""
    phys = PhysicsSettings(**{k: d[k] for k in asdict(PhysicsSettings()) if k in d})
    simd = {k: d[k] for k in asdict(SimulationSettings()) if k in d}
    if "method" in d:
        simd["sim_method"] = d["method"]
    return phys, SimulationSettings(**simd)


def load_settings(path: str):
    """Settings::load (settings/settings.cpp:14-59) through the library's parser."""
    s = SphSettings()
    check(lib().sph_load_settings(path.encode(), C.byref(s)))
    phys = PhysicsSettings(s.rest_density, s.k_viscosity, s.gravity, s.gas_stiffness, s.surface_tension)
    sim = SimulationSettings(s.time_step, s.max_vel, s.fluid_particle_radius, s.fluid_support_radius,
                             s.pcisph_max_iterations, s.pcisph_error_ratio, s.method, s.flags)
    return phys, sim


def _f4(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 2 and a.shape[1] == 3:
        a = np.concatenate([a, np.ones((a.shape[0], 1), np.float32)], axis=1)
    return np.ascontiguousarray(a.reshape(-1, 4))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class FluidSimulation:
    """FluidSimulation (fluidsimulation.h:23-92) backed by one libsph handle."""

    method = None

    def __init__(self, physics: PhysicsSettings, sim: SimulationSettings):
        if self.method is not None and sim.sim_method != self.method:
            sim = SimulationSettings(**{**asdict(sim), "sim_method": self.method})
        self._settings = make_settings(physics, sim)
        self._h = lib().sph_create(C.byref(self._settings))
        if not self._h:
            raise RuntimeError("sph_create failed: " + lib().sph_last_error().decode())

    @classmethod
    def _borrow(cls, handle, settings: SphSettings):
        """A view of a solver owned elsewhere (a loaded scene); close() leaves it."""
        obj = cls.__new__(cls)
        obj._settings = settings
        obj._h = handle
        obj._borrowed = True
        return obj

    def close(self):
        if getattr(self, "_h", None):
            if not getattr(self, "_borrowed", False):
                lib().sph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # -- reference interface -------------------------------------------------
    def simulate(self):
        check(lib().sph_step(self._h))

    def reset(self, physics: PhysicsSettings, sim: SimulationSettings):
        self._settings = make_settings(physics, sim)
        check(lib().sph_reset(self._h, C.byref(self._settings)))

    def particle_count(self) -> int:
        return check(lib().sph_particle_count(self._h))

    def add_boundary(self, samples, can_move=False, mass=0.0, R=None, com=None) -> int:
        """add_boundary with the body's surface_sampling() output and transform."""
        a = _f4(samples)
        R = None if R is None else np.ascontiguousarray(R, np.float32).reshape(9)
        com = None if com is None else np.ascontiguousarray(com, np.float32).reshape(3)
        return check(lib().sph_add_boundary(self._h, _ptr(a), a.shape[0], int(bool(can_move)), float(mass),
                                            _ptr(R), _ptr(com)))

    def set_rect_limits(self, width, height, depth):
        check(lib().sph_set_rect_limits(self._h, width, height, depth))

    def boundary_particle_count(self) -> int:
        return check(lib().sph_boundary_particle_count(self._h))

    def add_volume(self, particles):
        """add_volume with a FluidVolume whose particles() is this array."""
        a = _f4(particles)
        check(lib().sph_add_volume(self._h, _ptr(a), a.shape[0]))

    def add_box_volume(self, size, center):
        """add_volume(BoxVolume(size, center))."""
        s = np.ascontiguousarray(size, np.float32)
        c = np.ascontiguousarray(center, np.float32)
        check(lib().sph_add_box_volume(self._h, _ptr(s), _ptr(c)))

    # -- host-side state (replaces the GL VBO and Bullet reads) --------------
    def step(self, n: int = 1):
        check(lib().sph_step_n(self._h, n))

    def step_dt(self, dt: float):
        check(lib().sph_step_dt(self._h, dt))

    def synchronize(self):
        check(lib().sph_synchronize(self._h))

    def set_body_state(self, body, R=None, com=None, lin=None, ang=None):
        cv = lambda x, n: None if x is None else np.ascontiguousarray(x, np.float32).reshape(n)
        R, com, lin, ang = cv(R, 9), cv(com, 3), cv(lin, 3), cv(ang, 3)
        check(lib().sph_set_body_state(self._h, body, _ptr(R), _ptr(com), _ptr(lin), _ptr(ang)))

    def body_wrench(self, body):
        f = np.zeros(3, np.float32)
        t = np.zeros(3, np.float32)
        check(lib().sph_get_body_wrench(self._h, body, _ptr(f), _ptr(t)))
        return f, t

    def _get(self, fn):
        n = self.particle_count()
        out = np.zeros((n, 4), np.float32)
        check(fn(self._h, _ptr(out), n))
        return out

    def _set(self, fn, a):
        a = _f4(a)
        check(fn(self._h, _ptr(a), a.shape[0]))

    def positions(self):
        return self._get(lib().sph_get_positions)

    def velocities(self):
        return self._get(lib().sph_get_velocities)

    def half_velocities(self):
        return self._get(lib().sph_get_half_velocities)

    def set_positions(self, a):
        self._set(lib().sph_set_positions, a)

    def set_velocities(self, a):
        self._set(lib().sph_set_velocities, a)

    def set_half_velocities(self, a):
        self._set(lib().sph_set_half_velocities, a)

    def field(self, name: str) -> np.ndarray:
        fid = FIELDS[name]
        n = check(lib().sph_get_field(self._h, fid, None, 0))
        if name in ("ranges", "branges"):
            out = np.zeros((n, 3), np.int32)
        elif name in VEC4_FIELDS:
            out = np.zeros((n, 4), np.float32)
        elif name in U32_FIELDS:
            out = np.zeros(n, np.uint32)
        elif name in I32_FIELDS:
            out = np.zeros(n, np.int32)
        else:
            out = np.zeros(n, np.float32)
        check(lib().sph_get_field(self._h, fid, _ptr(out), n))
        return out

    def step_stats(self) -> dict:
        s = SphStepStats()
        check(lib().sph_get_step_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in SphStepStats._fields_}


    TUNED_KERNELS = ("normals", "forces_predict", "update_pressure", "pressure_force")

    def autotune(self, cache: str = None, trials: int = 3) -> bool:
        """Times the PCISPH list kernels' launch shapes and keeps the fastest
        (sph_autotune); True if the cache supplied them."""
        return check(lib().sph_autotune(self._h, None if cache is None else cache.encode(), int(trials))) == 1

    def launch(self, kernel) -> tuple:
        """(threads per block, grid) of a tuned kernel (name or SPH_TUNE_* id)."""
        k = self.TUNED_KERNELS.index(kernel) if isinstance(kernel, str) else int(kernel)
        b, g = C.c_int(), C.c_int()
        check(lib().sph_get_launch(self._h, k, C.byref(b), C.byref(g)))
        return b.value, g.value

    def set_launch(self, kernel, block: int, grid: int = 0):
        k = self.TUNED_KERNELS.index(kernel) if isinstance(kernel, str) else int(kernel)
        check(lib().sph_set_launch(self._h, k, int(block), int(grid)))

    def set_profiling(self, on=True, kernel: str = None):
        """HIP-event timing of every launch, or of one kernel's launches only."""
        check(lib().sph_set_profiling_filter(self._h, (kernel or "").encode()))
        check(lib().sph_set_profiling(self._h, int(on)))

    def kernel_times(self) -> dict:
        n = check(lib().sph_get_kernel_times(self._h, None, None, None, 0))
        names = C.create_string_buffer(48 * max(n, 1))
        launches = np.zeros(max(n, 1), np.int32)
        ms = np.zeros(max(n, 1), np.float64)
        check(lib().sph_get_kernel_times(self._h, names, _ptr(launches), _ptr(ms), n))
        out = {}
        for k in range(n):
            nm = names.raw[48 * k:48 * k + 48].split(b"\0", 1)[0].decode()
            out[nm] = (int(launches[k]), float(ms[k]))
        return out

    def reset_kernel_times(self):
        check(lib().sph_reset_kernel_times(self._h))


class PCISPHSimulation(FluidSimulation):
    """PCISPHSimulation (pcisphsimluation.h)."""
    method = PCISPH


class WCSPHSimulation(FluidSimulation):
    """WCSPHSimulation (wcsphsimluation.h)."""
    method = WCSPH


class FluidSimulationFactory:
    """FluidSimulationFactory::build_simulation (fluidsimulationfactory.h:13-40)."""

    @staticmethod
    def build_simulation(physics: PhysicsSettings, sim: SimulationSettings) -> FluidSimulation:
        if sim.sim_method == WCSPH:
            return WCSPHSimulation(physics, sim)
        if sim.sim_method == PCISPH:
            return PCISPHSimulation(physics, sim)
        raise RuntimeError("Unknown simulation method!")


def _sample(fn, *args) -> np.ndarray:
    n = check(fn(*args, None, 0))
    out = np.zeros((n, 4), np.float32)
    check(fn(*args, _ptr(out), n))
    return out


def box_volume_particles(size, center, radius) -> np.ndarray:
    """BoxVolume::particles (boxvolume.cpp:11-40)."""
    s = np.ascontiguousarray(size, np.float32)
    c = np.ascontiguousarray(center, np.float32)
    return _sample(lib().sph_sample_box_volume, _ptr(s), _ptr(c), float(radius))


def fishtank_sampling(width, height, depth, spacing) -> np.ndarray:
    """FishTank::surface_sampling (scene/fishtank.cpp:108-141)."""
    return _sample(lib().sph_sample_fishtank, float(width), float(height), float(depth), float(spacing))


def cube_sampling(side, spacing) -> np.ndarray:
    """Cube::surface_sampling (scene/cube.cpp:136-170)."""
    return _sample(lib().sph_sample_cube, float(side), float(spacing))


def wall_sampling(width, height, spacing) -> np.ndarray:
    """Wall::surface_sampling (scene/wall.cpp:142-152)."""
    return _sample(lib().sph_sample_wall, float(width), float(height), float(spacing))


def sphere_sampling(radius, spacing) -> np.ndarray:
    """Sphere::surface_sampling (scene/sphere.cpp:135-174), body-local."""
    return _sample(lib().sph_sample_sphere, float(radius), float(spacing))


def model_sampling(obj_path, spacing) -> np.ndarray:
    """Model::surface_sampling (scene/model.cpp:170-214): voxel corners of an
    OBJ mesh, body-local."""
    return _sample(lib().sph_sample_model, str(obj_path).encode(), float(spacing))


class LoadedScene:
    """Scene::load_scene (scene/scene.cpp:322-476) without rendering: the
    solver, the rigid world and the bodies of a scene file, advanced frame by
    frame as Scene::render does (scene.cpp:100-110)."""

    def __init__(self, path, physics: PhysicsSettings, sim: SimulationSettings, base_dir=None):
        from .rigid import RigidWorld
        self._settings = make_settings(physics, sim)
        L = lib()
        self._h = L.sph_scene_load(str(path).encode(), C.byref(self._settings),
                                   str(base_dir).encode() if base_dir else None)
        if not self._h:
            raise SphError(-5, L.sph_last_error().decode())
        cls = WCSPHSimulation if sim.sim_method == WCSPH else PCISPHSimulation
        self.simulation = cls._borrow(L.sph_scene_simulation(self._h), self._settings)
        self.world = RigidWorld._borrow(L.sph_scene_world(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self.simulation._h = None
            self.world._h = None
            lib().sph_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        check(lib().sph_scene_step(self._h, int(n)))

    def bodies(self) -> list:
        out = []
        for k in range(check(lib().sph_scene_body_count(self._h))):
            t, nm = C.create_string_buffer(16), C.create_string_buffer(64)
            sid, rid, smp = C.c_int(), C.c_int(), C.c_int()
            m = C.c_float()
            check(lib().sph_scene_body(self._h, k, t, nm, C.byref(sid), C.byref(rid), C.byref(m), C.byref(smp)))
            out.append({"type": t.value.decode(), "name": nm.value.decode(), "boundary_id": sid.value,
                        "rigid_id": rid.value, "mass": m.value, "samples": smp.value})
        return out


def load_scene(path, physics: PhysicsSettings, sim: SimulationSettings, base_dir=None) -> LoadedScene:
    return LoadedScene(path, physics, sim, base_dir)
