"""CPU tests of libsph's C ABI: it loads, exports every symbol include/sph.h
declares, and its host-only entry points (settings parser, samplers, error
reporting) behave like the reference. No GPU compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from fluids_amd import _lib, load_settings
from fluids_amd.simulation import box_volume_particles, cube_sampling, fishtank_sampling, wall_sampling

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sph.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 35
    for s in syms:
        assert hasattr(L, s), s
    # and the ctypes signatures cover them all
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_load_reference_configs():
    phys, sim = load_settings(os.path.join(GOLDEN, "config.pci"))
    assert sim.sim_method == _lib.PCISPH
    assert np.float32(sim.time_step) == np.float32(0.005)
    assert sim.pcisph_max_iterations == 3
    assert np.float32(phys.k_viscosity) == np.float32(1.7)
    assert np.float32(sim.fluid_support_radius) == np.float32(0.032)
    phys, sim = load_settings(os.path.join(GOLDEN, "config.wc"))
    assert sim.sim_method == _lib.WCSPH
    assert np.float32(phys.gas_stiffness) == np.float32(750)
    assert np.float32(phys.k_viscosity) == np.float32(0.00001)


def test_config_errors(tmp_path):
    """settings.cpp:37-39, 52-57 and fileparser.cpp:10-14, 33-36: a missing
    file, a missing key and an unknown method are errors with the
    reference's messages; unknown keys and '#' lines are ignored."""
    with pytest.raises(_lib.SphError, match="could not be parsed"):
        load_settings(str(tmp_path / "nope.cfg"))
    base = open(os.path.join(GOLDEN, "config.pci")).read()
    p = tmp_path / "a.cfg"
    p.write_text(base.replace("max_vel=20.0\n", ""))
    with pytest.raises(_lib.SphError, match="Option 'max_vel' was not found"):
        load_settings(str(p))
    p.write_text(base.replace("method=pcisph", "method=sph"))
    with pytest.raises(_lib.SphError, match="Unknown simulation method 'sph'"):
        load_settings(str(p))
    p.write_text(base + "\nfoo = 3\n  time_step   =  0.25   trailing\n")
    phys, sim = load_settings(str(p))
    assert np.float32(sim.time_step) == np.float32(0.25)


def test_last_error_is_per_call():
    L = _lib.lib()
    out = _lib.SphSettings()
    assert L.sph_load_settings(b"/nonexistent", C.byref(out)) == -5
    assert b"could not be parsed" in L.sph_last_error()
    assert L.sph_default_settings(C.byref(out)) == 0
    assert L.sph_last_error() == b""
    assert np.float32(out.rest_density) == np.float32(998.29)  # physicssettings.h:18


def test_flag_validation():
    """Unknown flag bits are rejected, and SPH_FAST_MATH is a single-scene
    option: a batch scene and the adjoint reject it before any device call."""
    L = _lib.lib()
    s = _lib.SphSettings()
    assert L.sph_default_settings(C.byref(s)) == 0
    s.method = _lib.WCSPH
    s.flags = _lib.FAST_MATH
    pos = np.zeros((1, 4), np.float32)
    p = pos.ctypes.data_as(C.POINTER(C.c_float))
    assert not L.sph_adjoint_create(C.byref(s), 1.0, 1.0, 1.0, p, p, p, 1, 1)
    assert b"adjoint: SPH_FAST_MATH is not supported" in L.sph_last_error()
    assert L.sph_batch_add_scene(None, C.byref(s), 1.0, 1.0, 1.0, p, 1) < 0
    assert b"batch: SPH_FAST_MATH is not supported" in L.sph_last_error()
    s.method = _lib.PCISPH
    s.flags = 8
    assert not L.sph_create(C.byref(s))
    assert b"unknown flag bits" in L.sph_last_error()


def test_samplers_match_oracle():
    assert np.array_equal(fishtank_sampling(2, 1, 1, 0.02), oracle.fishtank(2, 1, 1, 0.02))
    assert np.array_equal(fishtank_sampling(2, 1, 1, 0.005), oracle.fishtank(2, 1, 1, 0.005))
    assert np.array_equal(cube_sampling(0.3, 0.01), oracle.cube(0.3, 0.01))
    assert np.array_equal(wall_sampling(0.7, 0.3, 0.02), oracle.wall(0.7, 0.3, 0.02))
    for size, center, r in [([0.5, 0.64, 0.8], [0.73, -0.16, 0], 0.01), ([0.5, 1.0, 0.7], [0.74, 0, -0.1], 0.005)]:
        a = box_volume_particles(size, center, r)
        assert np.array_equal(a, oracle.box_volume(size, center, r))
    a = box_volume_particles([0.5, 1.0, 0.7], [0.74, 0, -0.1], 0.005)
    assert a.shape[0] == 50 * 100 * 70
    # lattice order y, x, z (boxvolume.cpp:63-65)
    assert a[1, 2] > a[0, 2] and a[70, 0] > a[0, 0] and a[70 * 50, 1] > a[0, 1]


def test_sampler_counts():
    # the single-layer tanks of the benchmark scenes
    assert fishtank_sampling(2, 1, 1, 0.005).shape[0] == 321401
    assert cube_sampling(0.3, 0.01).shape[0] == 5280
