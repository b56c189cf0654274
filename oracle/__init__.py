"""ctypes wrapper of the CPU oracle — TEST INFRASTRUCTURE ONLY.

The oracle restates the reference SPH step on the CPU (see sph_oracle.cpp for
the pinned arithmetic and the reference citations). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
checker; the product never imports it.
"""
import ctypes as C
import os

import numpy as np

from . import build as _build

_LIB = None

# Field ids (same meaning as the library's SPH_FIELD_* in include/sph.h).
FIELDS = [
    "keys", "perm", "ncount", "nlist", "bncount", "bnlist",
    "pos_sorted", "vel_sorted", "density", "pressure", "err", "normal",
    "force_other", "force_pressure", "pos_pred", "accel",
    "pos", "vel", "vel_half", "bpos_sorted", "bphi_sorted", "bkeys", "bperm",
    "bphi", "bforce", "btorque", "ranges", "branges", "vel_half_sorted",
]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
VEC4 = {"pos_sorted", "vel_sorted", "normal", "force_other", "force_pressure",
        "pos_pred", "accel", "pos", "vel", "vel_half", "bpos_sorted", "bforce",
        "btorque", "vel_half_sorted"}
INTS = {"keys", "perm", "ncount", "nlist", "bncount", "bnlist", "bkeys", "bperm"}


class Settings(C.Structure):
    _fields_ = [
        ("time_step", C.c_float), ("max_vel", C.c_float),
        ("fluid_particle_radius", C.c_float), ("fluid_support_radius", C.c_float),
        ("pcisph_max_iterations", C.c_int), ("pcisph_error_ratio", C.c_float),
        ("method", C.c_int),
        ("rest_density", C.c_float), ("k_viscosity", C.c_float), ("gravity", C.c_float),
        ("gas_stiffness", C.c_float), ("surface_tension", C.c_float),
        ("flags", C.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        # SPH_ORACLE_LIB names another build of the same source (the mutation
        # check of tests/test_oracle_mutants.py builds altered copies)
        path = os.environ.get("SPH_ORACLE_LIB") or _build.LIB
        if path == _build.LIB and not os.path.exists(path):
            _build.build()
        L = C.CDLL(path)
        P, I, F = C.c_void_p, C.c_int, C.c_float
        L.oracle_create.restype = P
        L.oracle_create.argtypes = [C.POINTER(Settings)]
        for name, args in [
            ("oracle_destroy", [P]), ("oracle_set_twin", [P, I]), ("oracle_set_threads", [I]),
            ("oracle_set_rect_limits", [P, F, F, F]), ("oracle_add_volume", [P, P, I]),
            ("oracle_set_body_state", [P, I, P, P, P, P]),
            ("oracle_get_body_wrench", [P, I, P, P, P, P]),
            ("oracle_set_state", [P, P, P, P, I]), ("oracle_step", [P, I]),
        ]:
            getattr(L, name).restype = None
            getattr(L, name).argtypes = args
        L.oracle_add_boundary.restype = I
        L.oracle_add_boundary.argtypes = [P, P, I, I, F, P, P]
        for name in ["oracle_particle_count", "oracle_boundary_count", "oracle_iterations",
                     "oracle_grid_cps", "oracle_max_threads"]:
            getattr(L, name).restype = I
        L.oracle_particle_count.argtypes = [P]
        L.oracle_boundary_count.argtypes = [P]
        L.oracle_iterations.argtypes = [P]
        L.oracle_grid_cps.argtypes = [P]
        L.oracle_max_threads.argtypes = []
        L.oracle_edge_reads.restype = C.c_longlong
        L.oracle_edge_reads.argtypes = [P]
        L.oracle_delta.restype = F
        L.oracle_delta.argtypes = [P]
        L.oracle_constants.restype = None
        L.oracle_constants.argtypes = [P, P, I]
        L.oracle_get_field.restype = I
        L.oracle_get_field.argtypes = [P, I, P, I]
        L.oracle_morton3.restype = C.c_uint32
        L.oracle_morton3.argtypes = [C.c_uint32] * 3
        L.oracle_define_literal.restype = F
        L.oracle_define_literal.argtypes = [F]
        L.oracle_deduce_delta.restype = F
        L.oracle_deduce_delta.argtypes = [F, F, F, F, F]
        L.oracle_boundary_phi.restype = None
        L.oracle_boundary_phi.argtypes = [P, I, F, F, I, P]
        L.oracle_box_volume.restype = I
        L.oracle_box_volume.argtypes = [P, P, F, P, I]
        L.oracle_fishtank.restype = I
        L.oracle_fishtank.argtypes = [F, F, F, F, P, I]
        L.oracle_cube.restype = I
        L.oracle_cube.argtypes = [F, F, P, I]
        L.oracle_wall.restype = I
        L.oracle_wall.argtypes = [F, F, F, P, I]
        D = C.c_double
        L.oracle_wcsph_rollout64.restype = I
        L.oracle_wcsph_rollout64.argtypes = [C.POINTER(Settings), D, D, D, P, P, P, I, I, D, D, D, P, P, P, P]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _f4(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 2 and a.shape[1] == 3:
        a = np.concatenate([a, np.ones((a.shape[0], 1), np.float32)], axis=1)
    return np.ascontiguousarray(a.reshape(-1, 4))


CONSTANT_NAMES = [
    "pci_mass", "pci_mass_d", "pci_h_d", "pci_rho0_d", "pci_st_main", "pci_st_term",
    "pci_default_eval", "pci_default_grad", "pci_visc_lapl", "pci_pressure_grad", "pci_delta",
    "wc_mass", "wc_poly6_eval", "wc_poly6_grad", "wc_spiky_grad", "wc_visc_lapl",
    "wc_st_main", "wc_st_term", "wc_mass_d", "wc_h_d",
    "b_h_d", "b_poly6_eval", "b_spiky_grad", "grid_sqr_support", "grid_cps",
]


class Oracle:
    """Stateful CPU restatement of PCISPHSimulation / WCSPHSimulation."""

    def __init__(self, settings: dict, twin: bool = False):
        s = Settings(**{k: v for k, v in settings.items() if k in dict(Settings._fields_)})
        self._s = s
        self.h = lib().oracle_create(C.byref(s))
        if twin:
            lib().oracle_set_twin(self.h, 1)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().oracle_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_rect_limits(self, w, h, d):
        lib().oracle_set_rect_limits(self.h, w, h, d)

    def add_volume(self, xyzw):
        a = _f4(xyzw)
        lib().oracle_add_volume(self.h, _ptr(a), a.shape[0])

    def add_boundary(self, xyzw, can_move=False, mass=0.0, R=None, com=None):
        a = _f4(xyzw)
        R = None if R is None else np.ascontiguousarray(R, np.float32).reshape(9)
        com = None if com is None else np.ascontiguousarray(com, np.float32).reshape(3)
        return lib().oracle_add_boundary(self.h, _ptr(a), a.shape[0], int(can_move), float(mass),
                                         _ptr(R), _ptr(com))

    def set_body_state(self, body, R=None, com=None, lin=None, ang=None):
        cv = lambda x, n: None if x is None else np.ascontiguousarray(x, np.float32).reshape(n)
        R, com, lin, ang = cv(R, 9), cv(com, 3), cv(lin, 3), cv(ang, 3)
        lib().oracle_set_body_state(self.h, body, _ptr(R), _ptr(com), _ptr(lin), _ptr(ang))

    def body_wrench(self, body):
        f = np.zeros(3, np.float32)
        t = np.zeros(3, np.float32)
        f64 = np.zeros(3, np.float64)
        t64 = np.zeros(3, np.float64)
        lib().oracle_get_body_wrench(self.h, body, _ptr(f), _ptr(t), _ptr(f64), _ptr(t64))
        return f, t, f64, t64

    def set_state(self, pos, vel=None, vel_half=None):
        pos = _f4(pos)
        vel = None if vel is None else _f4(vel)
        vh = None if vel_half is None else _f4(vel_half)
        lib().oracle_set_state(self.h, _ptr(pos), _ptr(vel), _ptr(vh), pos.shape[0])

    def step(self, n=1):
        lib().oracle_step(self.h, n)

    @property
    def particle_count(self):
        return lib().oracle_particle_count(self.h)

    @property
    def boundary_count(self):
        return lib().oracle_boundary_count(self.h)

    @property
    def iterations(self):
        return lib().oracle_iterations(self.h)

    @property
    def edge_reads(self):
        return lib().oracle_edge_reads(self.h)

    @property
    def cps(self):
        return lib().oracle_grid_cps(self.h)

    def constants(self):
        out = np.zeros(len(CONSTANT_NAMES), np.float32)
        lib().oracle_constants(self.h, _ptr(out), len(CONSTANT_NAMES))
        return dict(zip(CONSTANT_NAMES, out.tolist()))

    def field(self, name, twin=False):
        fid = FIELD_ID[name]
        n = lib().oracle_get_field(self.h, fid, None, int(twin))
        if n < 0:
            raise KeyError(name)
        if twin:
            if name in VEC4:
                out = np.zeros((n, 3), np.float64)
            else:
                out = np.zeros(n, np.float64)
        elif name in ("ranges", "branges"):
            out = np.zeros((n, 3), np.int32)
        elif name in VEC4:
            out = np.zeros((n, 4), np.float32)
        elif name in INTS:
            out = np.zeros(n, np.uint32 if name in ("keys", "bkeys") else np.int32)
        else:
            out = np.zeros(n, np.float32)
        lib().oracle_get_field(self.h, fid, _ptr(out), int(twin))
        return out


def set_threads(n: int):
    lib().oracle_set_threads(int(n))


def morton3(x, y, z):
    return lib().oracle_morton3(x, y, z)


def define_literal(v):
    return lib().oracle_define_literal(v)


def deduce_delta(h, r, dt, mass, rho0):
    return lib().oracle_deduce_delta(h, r, dt, mass, rho0)


def boundary_phi(xyzw, h, rho0, use_grid=False):
    a = _f4(xyzw)
    phi = np.zeros(a.shape[0], np.float32)
    lib().oracle_boundary_phi(_ptr(a), a.shape[0], h, rho0, int(use_grid), _ptr(phi))
    return phi


def _sample(fn, *args):
    n = fn(*args, None, 0)
    out = np.zeros((n, 4), np.float32)
    fn(*args, _ptr(out), n)
    return out


def box_volume(size, center, radius):
    s = np.ascontiguousarray(size, np.float32)
    c = np.ascontiguousarray(center, np.float32)
    return _sample(lib().oracle_box_volume, _ptr(s), _ptr(c), float(radius))


def fishtank(w, h, d, spacing):
    return _sample(lib().oracle_fishtank, float(w), float(h), float(d), float(spacing))


def cube(side, spacing):
    return _sample(lib().oracle_cube, float(side), float(spacing))


def wall(w, h, spacing):
    return _sample(lib().oracle_wall, float(w), float(h), float(spacing))


def wcsph_rollout64(settings: dict, container, pos0, vel0=None, vh0=None, steps=1, k_visc=None, k_gas=None,
                    sigma=None):
    """WCSPH in double precision with its own neighbour search (for finite
    differences): returns (pos, vel, vel_half, max list length), (n, 3)
    float64 arrays in the input's particle order."""
    st = Settings(**{k: v for k, v in settings.items() if k in dict(Settings._fields_)})
    x = np.ascontiguousarray(np.asarray(pos0, np.float64)[:, :3])
    n = x.shape[0]
    v = None if vel0 is None else np.ascontiguousarray(np.asarray(vel0, np.float64)[:, :3])
    w = None if vh0 is None else np.ascontiguousarray(np.asarray(vh0, np.float64)[:, :3])
    po, vo, wo = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    mc = C.c_int(0)
    kv = settings["k_viscosity"] if k_visc is None else k_visc
    kg = settings["gas_stiffness"] if k_gas is None else k_gas
    sg = settings["surface_tension"] if sigma is None else sigma
    lib().oracle_wcsph_rollout64(C.byref(st), float(container[0]), float(container[1]), float(container[2]),
                                 _ptr(x), _ptr(v), _ptr(w), n, int(steps), float(kv), float(kg), float(sg),
                                 _ptr(po), _ptr(vo), _ptr(wo), C.cast(C.byref(mc), C.c_void_p))
    return po, vo, wo, mc.value
