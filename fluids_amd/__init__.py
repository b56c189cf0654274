"""MI355X (gfx950) SPH solver: a drop-in for the OpenCL simulation path of
paivett/fluids, behind the C ABI of include/sph.h (libsph.so)."""
from ._lib import FAST_MATH, FIX_WCSPH_DENSITY, PCISPH, WCSPH, SphError
from .simulation import (FluidSimulation, FluidSimulationFactory, PCISPHSimulation, PhysicsSettings,
                         SimulationSettings, WCSPHSimulation, box_volume_particles, cube_sampling,
                         fishtank_sampling, load_scene, load_settings, model_sampling, settings_from_dict,
                         sphere_sampling, wall_sampling)

__all__ = [
    "FluidSimulation", "PCISPHSimulation", "WCSPHSimulation", "FluidSimulationFactory",
    "PhysicsSettings", "SimulationSettings", "load_settings", "settings_from_dict",
    "box_volume_particles", "fishtank_sampling", "cube_sampling", "wall_sampling", "sphere_sampling",
    "model_sampling", "load_scene",
    "SphError", "WCSPH", "PCISPH", "FIX_WCSPH_DENSITY", "FAST_MATH",
]
