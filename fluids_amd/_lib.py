"""ctypes binding of libsph.so (the C ABI of include/sph.h).

The library is loaded from this package directory. There is no fallback: if
it is missing or fails to load, every entry point raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SPH_LIB names another build of the library (the host-sanitizer build of
# tests/test_sanitizers.py); by default the in-tree libsph.so
LIB_PATH = os.environ.get("SPH_LIB") or os.path.join(HERE, "libsph.so")

SPH_OK = 0
WCSPH, PCISPH = 0, 1
FIX_WCSPH_DENSITY = 1
FAST_MATH = 2  # single-scene kernels with approximate division and square root (SPH_FAST_MATH)
LIST_CAP = 50

FIELDS = {
    "keys": 0, "perm": 1, "ncount": 2, "nlist": 3, "bncount": 4, "bnlist": 5,
    "pos_sorted": 6, "vel_sorted": 7, "density": 8, "pressure": 9, "err": 10, "normal": 11,
    "force_other": 12, "force_pressure": 13, "pos_pred": 14, "accel": 15,
    "pos": 16, "vel": 17, "vel_half": 18, "bpos_sorted": 19, "bphi_sorted": 20, "bkeys": 21,
    "bperm": 22, "bphi": 23, "bforce": 24, "btorque": 25, "ranges": 26, "branges": 27,
    "vel_half_sorted": 28,
}
VEC4_FIELDS = {"pos_sorted", "vel_sorted", "normal", "force_other", "force_pressure", "pos_pred",
               "accel", "pos", "vel", "vel_half", "bpos_sorted", "bforce", "btorque", "vel_half_sorted"}
U32_FIELDS = {"keys", "bkeys"}
I32_FIELDS = {"perm", "ncount", "nlist", "bncount", "bnlist", "bperm"}


class SphSettings(C.Structure):
    _fields_ = [
        ("time_step", C.c_float), ("max_vel", C.c_float),
        ("fluid_particle_radius", C.c_float), ("fluid_support_radius", C.c_float),
        ("pcisph_max_iterations", C.c_int), ("pcisph_error_ratio", C.c_float),
        ("method", C.c_int),
        ("rest_density", C.c_float), ("k_viscosity", C.c_float), ("gravity", C.c_float),
        ("gas_stiffness", C.c_float), ("surface_tension", C.c_float),
        ("flags", C.c_int),
    ]


class SphStepStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("fluid_list_overflows", C.c_int),
        ("boundary_list_overflows", C.c_int), ("edge_reads", C.c_int),
        ("out_of_grid", C.c_int), ("nan_positions", C.c_int),
        ("fluid_neighbours", C.c_longlong), ("boundary_neighbours", C.c_longlong),
        ("max_density_error", C.c_float),
    ]


# name -> (restype, argtypes); every symbol include/sph.h declares
P, I, F, LL = C.c_void_p, C.c_int, C.c_float, C.c_longlong
FP = C.POINTER(C.c_float)
SIGNATURES = {
    "sph_last_error": (C.c_char_p, []),
    "sph_version": (C.c_char_p, []),
    "sph_default_settings": (I, [C.POINTER(SphSettings)]),
    "sph_load_settings": (I, [C.c_char_p, C.POINTER(SphSettings)]),
    "sph_create": (P, [C.POINTER(SphSettings)]),
    "sph_destroy": (I, [P]),
    "sph_reset": (I, [P, C.POINTER(SphSettings)]),
    "sph_set_rect_limits": (I, [P, F, F, F]),
    "sph_add_volume": (I, [P, P, I]),
    "sph_add_box_volume": (I, [P, P, P]),
    "sph_add_boundary": (I, [P, P, I, I, F, P, P]),
    "sph_set_body_state": (I, [P, I, P, P, P, P]),
    "sph_get_body_wrench": (I, [P, I, P, P]),
    "sph_step": (I, [P]),
    "sph_step_n": (I, [P, I]),
    "sph_step_dt": (I, [P, F]),
    "sph_synchronize": (I, [P]),
    "sph_particle_count": (I, [P]),
    "sph_boundary_particle_count": (I, [P]),
    "sph_get_positions": (I, [P, P, I]),
    "sph_set_positions": (I, [P, P, I]),
    "sph_get_velocities": (I, [P, P, I]),
    "sph_set_velocities": (I, [P, P, I]),
    "sph_get_half_velocities": (I, [P, P, I]),
    "sph_set_half_velocities": (I, [P, P, I]),
    "sph_get_field": (I, [P, I, P, LL]),
    "sph_get_step_stats": (I, [P, C.POINTER(SphStepStats)]),
    "sph_device_state": (I, [P, P, P, P]),
    "sph_stream": (P, [P]),
    "sph_autotune": (I, [P, C.c_char_p, I]),
    "sph_get_launch": (I, [P, I, P, P]),
    "sph_set_launch": (I, [P, I, I, I]),
    "sph_sort_pairs": (I, [P, P, I, I, P]),
    "sph_sort_pairs_noalloc": (I, [P, P, I, I, P]),
    "sph_sort_near": (I, [P, P, P, P, I, P]),
    "sph_set_profiling": (I, [P, I]),
    "sph_set_profiling_filter": (I, [P, C.c_char_p]),
    "sph_get_kernel_times": (I, [P, P, P, P, I]),
    "sph_reset_kernel_times": (I, [P]),
    "sph_sample_box_volume": (I, [P, P, F, P, I]),
    "sph_sample_fishtank": (I, [F, F, F, F, P, I]),
    "sph_sample_cube": (I, [F, F, P, I]),
    "sph_sample_wall": (I, [F, F, F, P, I]),
    "sph_sample_sphere": (I, [F, F, P, I]),
    "sph_sample_model": (I, [C.c_char_p, F, P, I]),
    "sph_scene_load": (P, [C.c_char_p, C.POINTER(SphSettings), C.c_char_p]),
    "sph_scene_destroy": (I, [P]),
    "sph_scene_simulation": (P, [P]),
    "sph_scene_world": (P, [P]),
    "sph_scene_step": (I, [P, I]),
    "sph_scene_body_count": (I, [P]),
    "sph_scene_body": (I, [P, I, P, P, P, P, P, P]),
    "sph_frames_open": (P, [P, C.c_char_p, I]),
    "sph_frames_write": (I, [P]),
    "sph_frames_close": (I, [P]),
    "sph_perf_open": (P, [P, C.c_char_p]),
    "sph_perf_frame": (I, [P]),
    "sph_perf_close": (I, [P]),
    "sph_rigid_create": (P, [F]),
    "sph_rigid_destroy": (I, [P]),
    "sph_rigid_set_container": (I, [P, F, F, F, P]),
    "sph_rigid_add_box": (I, [P, F, P, P, P]),
    "sph_rigid_add_sphere": (I, [P, F, F, P]),
    "sph_rigid_add_static": (I, [P, P, P]),
    "sph_rigid_apply_force_n_torque": (I, [P, I, P, P]),
    "sph_rigid_step": (I, [P, F]),
    "sph_rigid_get_state": (I, [P, I, P, P, P, P, P]),
    "sph_rigid_set_state": (I, [P, I, P, P, P, P]),
    "sph_rigid_body_count": (I, [P]),
    "sph_coupled_step": (I, [P, P, I, P, P]),
    "sph_batch_create": (P, []),
    "sph_batch_destroy": (I, [P]),
    "sph_batch_add_scene": (I, [P, C.POINTER(SphSettings), F, F, F, P, I]),
    "sph_batch_finalize": (I, [P]),
    "sph_batch_autotune": (I, [P, C.c_char_p, I]),
    "sph_batch_get_launch": (I, [P, I, P, P, P]),
    "sph_batch_set_launch": (I, [P, I, I, I, I]),
    "sph_batch_set_graph": (I, [P, I]),
    "sph_batch_step": (I, [P, I]),
    "sph_batch_synchronize": (I, [P]),
    "sph_batch_scene_count": (I, [P]),
    "sph_batch_particle_count": (I, [P, I]),
    "sph_batch_total_particles": (I, [P]),
    "sph_batch_get_positions": (I, [P, I, P, I]),
    "sph_batch_get_velocities": (I, [P, I, P, I]),
    "sph_batch_get_half_velocities": (I, [P, I, P, I]),
    "sph_batch_set_positions": (I, [P, I, P, I]),
    "sph_batch_set_velocities": (I, [P, I, P, I]),
    "sph_batch_get_density": (I, [P, I, P, I]),
    "sph_batch_get_neighbour_counts": (I, [P, I, P, I]),
    "sph_batch_set_profiling": (I, [P, I, C.c_char_p]),
    "sph_batch_get_kernel_times": (I, [P, P, P, P, I]),
    "sph_adjoint_create": (P, [C.POINTER(SphSettings), F, F, F, P, P, P, I, I]),
    "sph_adjoint_destroy": (I, [P]),
    "sph_adjoint_set_initial": (I, [P, P, P, P]),
    "sph_adjoint_set_params": (I, [P, F, F, F]),
    "sph_adjoint_forward": (I, [P, I]),
    "sph_adjoint_final_state": (I, [P, P, P, P]),
    "sph_adjoint_backward": (I, [P, P, P, P, P, P, P, P]),
    "sph_adjoint_timings": (I, [P, P, P]),
}

_LIB = None


class SphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libsph error {code}: {msg}")
        self.code = code


def lib():
    """Loads libsph.so; raises if it is missing (no fallback path exists)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m fluids_amd.build` "
                "(or __graft_entry__.build()); the solver has no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc < 0:
        raise SphError(rc, lib().sph_last_error().decode())
    return rc
