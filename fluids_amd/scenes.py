"""Seeded synthetic scenes for the benchmark configurations.

All randomness comes from SplitMix64(seed): u64 -> u = (u64 >> 40) * 2^-24,
identical in C++ and Python. Fluid lattices come from BoxVolume::particles
(boxvolume.cpp:11-40) and are jittered per particle, in lattice order, by
(2u - 1) * 1e-3 * spacing per axis (x, y, z draws per particle), evaluated in
float64 and added to the float32 position in float32. Walls are
FishTank::surface_sampling (scene/fishtank.cpp:108-141) and the rigid cube is
Cube::surface_sampling (scene/cube.cpp:136-170).

Placement rule: fluid lattice points stay at least one spacing away from any
sampled wall; a fluid particle on a boundary particle gives r = 0, which the
reference's boundary pressure terms divide by (pcisph.cl:493-498).
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from ._lib import FIX_WCSPH_DENSITY, PCISPH, WCSPH
from .simulation import box_volume_particles, cube_sampling, fishtank_sampling

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Outputs offset+1 .. offset+count of SplitMix64 seeded with `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int, offset: int = 0) -> np.ndarray:
    return (splitmix64(seed, count, offset) >> np.uint64(40)).astype(np.float64) * 2.0 ** -24


class Rng:
    """Sequential SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = seed
        self.pos = 0

    def uniform(self, n: int = 1) -> np.ndarray:
        u = uniform01(self.seed, n, self.pos)
        self.pos += n
        return u

    def u64(self) -> int:
        v = int(splitmix64(self.seed, 1, self.pos)[0])
        self.pos += 1
        return v


def jitter(p: np.ndarray, spacing: float, rng: Rng) -> np.ndarray:
    u = rng.uniform(3 * p.shape[0]).reshape(-1, 3)
    j = ((2.0 * u - 1.0) * 1e-3 * spacing).astype(np.float32)
    out = p.copy()
    out[:, :3] = (p[:, :3] + j).astype(np.float32)
    out[:, 3] = 1.0
    return out


def quat_to_rows(q) -> np.ndarray:
    """btMatrix3x3::setRotation for a quaternion (x, y, z, w), normalised by
    its squared length like Bullet; returns the basis rows (row-major)."""
    x, y, z, w = [float(v) for v in q]
    d = x * x + y * y + z * z + w * w
    s = 2.0 / d
    xs, ys, zs = x * s, y * s, z * s
    wx, wy, wz = w * xs, w * ys, w * zs
    xx, xy, xz = x * xs, x * ys, x * zs
    yy, yz, zz = y * ys, y * zs, z * zs
    R = np.array([[1.0 - (yy + zz), xy - wz, xz + wy],
                  [xy + wz, 1.0 - (xx + zz), yz - wx],
                  [xz - wy, yz + wx, 1.0 - (xx + yy)]], np.float32)
    return R.reshape(9)


@dataclass
class BodySpec:
    samples: np.ndarray
    mass: float
    can_move: bool
    R: np.ndarray
    com: np.ndarray
    name: str = ""


@dataclass
class Scene:
    name: str
    settings: dict
    container: tuple
    volumes: List[np.ndarray]
    bodies: List[BodySpec] = field(default_factory=list)
    steps: int = 1
    notes: str = ""

    @property
    def n_fluid(self) -> int:
        return int(sum(v.shape[0] for v in self.volumes))

    @property
    def n_boundary(self) -> int:
        return int(sum(b.samples.shape[0] for b in self.bodies))


def _box(size, center, radius, rng: Optional[Rng]):
    p = box_volume_particles(size, center, radius)
    return jitter(p, 2 * radius, rng) if rng is not None else p


def _tank_body(w, h, d, spacing) -> BodySpec:
    s = fishtank_sampling(w, h, d, spacing)
    return BodySpec(s, 0.0, False, np.eye(3, dtype=np.float32).reshape(9), np.zeros(3, np.float32), "tank")


def c1_wcsph_dambreak(seed: int = 101, fix_density: bool = True) -> Scene:
    """WCSPH dam break: 25x32x40 = 32 000 particles on a 0.02 m lattice in a
    sampled 2x1x1 m tank. Tait EOS with gamma 7, c0 = 20 m/s:
    gas_stiffness = rho0 c0^2 / 7. Monaghan alpha = 0.02 maps onto the
    reference's Laplacian viscosity as k_viscosity = nu * m with
    nu = alpha * (h/2) * c0 / 10 (the WCSPH viscosity term carries no mass,
    wcsph.cl:157,167). surface_tension and max_vel from config.wc."""
    r, h, rho0, c0, alpha = 0.01, 0.04, 1000.0, 20.0, 0.02
    m = rho0 * (2 * r) ** 3
    nu = alpha * (h / 2) * c0 / 10
    settings = dict(time_step=2e-4, max_vel=10.0, fluid_particle_radius=r, fluid_support_radius=h,
                    pcisph_max_iterations=7, pcisph_error_ratio=0.01, method=WCSPH, rest_density=rho0,
                    k_viscosity=nu * m, gravity=9.81, gas_stiffness=rho0 * c0 * c0 / 7.0,
                    surface_tension=0.001, flags=FIX_WCSPH_DENSITY if fix_density else 0)
    rng = Rng(seed)
    vol = _box([0.5, 0.64, 0.8], [0.73, -0.16, 0.0], r, rng)
    return Scene("wcsph_dambreak_32k", settings, (2.0, 1.0, 1.0), [vol], [_tank_body(2.0, 1.0, 1.0, 2 * r)],
                 steps=1000)


def c2_pcisph_double_dambreak(seed: int = 202, spacing: float = 0.01, with_cube: bool = True) -> Scene:
    """PCISPH double dam break (data/scenes/double_dambreak.json) with the
    0.3 m, 12 kg cube of data/scenes/many_objects.json:31-37; tank walls are
    the integrators' clamp box as in the reference (scene.cpp:376-382)."""
    r = spacing / 2
    settings = dict(time_step=1e-3, max_vel=20.0, fluid_particle_radius=r, fluid_support_radius=2 * spacing,
                    pcisph_max_iterations=7, pcisph_error_ratio=0.01, method=PCISPH, rest_density=1000.0,
                    k_viscosity=1.7, gravity=9.8, gas_stiffness=500.0, surface_tension=1.0, flags=0)
    rng = Rng(seed)
    v1 = _box([0.5, 1.0, 0.7], [0.74, 0.0, -0.1], r, rng)
    v2 = _box([0.5, 1.0, 0.7], [-0.74, 0.0, 0.1], r, rng)
    bodies = []
    if with_cube:
        bodies.append(BodySpec(cube_sampling(0.3, spacing), 12.0, True, quat_to_rows([0, 1, 0, 1]),
                               np.array([0.0, -0.35, 0.0], np.float32), "cube"))
    return Scene("pcisph_double_dambreak", settings, (2.0, 1.0, 1.0), [v1, v2], bodies, steps=500)


def _c3_dims(n: int):
    # 125k and 500k: one GPU's share of the 1M and 4M scenes over 8 GPUs
    table = {125_000: (250, 25, 20), 500_000: (250, 40, 50), 1_000_000: (250, 50, 80), 4_000_000: (250, 100, 160)}
    if n in table:
        return table[n]
    # generic: a slab 250 wide and 5:8 in y:z
    nx = 250
    rest = n // nx
    ny = max(1, int(round((rest * 5 / 8) ** 0.5)))
    nz = max(1, rest // ny)
    return nx, ny, nz


def c3_pcisph_throughput(n: int = 1_000_000, seed: int = 303, tank: bool = True) -> Scene:
    """PCISPH throughput scene: exactly n lattice points at 0.005 m (h = 0.01)
    split into k in {1, 2, 4} boxes (from the seed) placed without overlap on
    the floor of a 2x1x1 m tank, one spacing above the floor and away from the
    walls, plus the tank sampled as one layer of boundary particles."""
    spacing, r = 0.005, 0.0025
    settings = dict(time_step=1e-3, max_vel=20.0, fluid_particle_radius=r, fluid_support_radius=2 * spacing,
                    pcisph_max_iterations=7, pcisph_error_ratio=0.01, method=PCISPH, rest_density=1000.0,
                    k_viscosity=1.7, gravity=9.8, gas_stiffness=500.0, surface_tension=1.0, flags=0)
    rng = Rng(seed)
    nx, ny, nz = _c3_dims(n)
    k = [1, 2, 4][rng.u64() % 3]
    if k == 2:
        dims = [(nx // 2, ny, nz)] * 2
    elif k == 4:
        dims = [(nx // 2, ny, nz // 2)] * 4
    else:
        dims = [(nx, ny, nz)]
    placed = []
    W, H, D = 2.0, 1.0, 1.0
    margin = 2 * spacing
    vols = []
    for (bx, by, bz) in dims:
        # size slightly under bx*spacing so that ceil(size/spacing) == bx
        size = np.array([bx, by, bz], np.float64) * spacing * (1 - 1e-5)
        for _ in range(10000):
            u = rng.uniform(2)
            lo_x = -W / 2 + margin + u[0] * (W - 2 * margin - size[0])
            lo_z = -D / 2 + margin + u[1] * (D - 2 * margin - size[2])
            box = (lo_x, lo_z, lo_x + size[0], lo_z + size[2])
            if all(box[2] + margin < b[0] or b[2] + margin < box[0] or box[3] + margin < b[1] or
                   b[3] + margin < box[1] for b in placed):
                placed.append(box)
                break
        else:
            raise RuntimeError("could not place the fluid boxes")
        center = [lo_x + size[0] / 2, -H / 2 + spacing + size[1] / 2, lo_z + size[2] / 2]
        v = _box(size.astype(np.float32), np.array(center, np.float32), r, rng)
        assert v.shape[0] == bx * by * bz, (v.shape, bx, by, bz)
        vols.append(v)
    bodies = [_tank_body(W, H, D, spacing)] if tank else []
    sc = Scene(f"pcisph_throughput_{n // 1000}k", settings, (W, H, D), vols, bodies, steps=200)
    assert sc.n_fluid == n
    return sc


def small_pcisph(seed: int = 7, n_side=(12, 10, 8), spacing=0.02, tank=True, cube=False) -> Scene:
    """A few-thousand-particle PCISPH scene for parity tests."""
    r = spacing / 2
    settings = dict(time_step=1e-3, max_vel=20.0, fluid_particle_radius=r, fluid_support_radius=2 * spacing,
                    pcisph_max_iterations=7, pcisph_error_ratio=0.01, method=PCISPH, rest_density=1000.0,
                    k_viscosity=1.7, gravity=9.8, gas_stiffness=500.0, surface_tension=1.0, flags=0)
    rng = Rng(seed)
    size = np.array(n_side, np.float64) * spacing * (1 - 1e-5)
    W, H, D = 1.2, 0.8, 0.6
    # one spacing above the tank floor (y = -H/2), away from the walls
    center = [0.3, -H / 2 + spacing + size[1] / 2, 0.05]
    vol = _box(size.astype(np.float32), np.array(center, np.float32), r, rng)
    bodies = [_tank_body(W, H, D, spacing)] if tank else []
    if cube:
        bodies.append(BodySpec(cube_sampling(0.12, spacing), 2.0, True, quat_to_rows([0, 0.3, 0, 1]),
                               np.array([-0.25, -0.25, 0.0], np.float32), "cube"))
    return Scene("small_pcisph", settings, (W, H, D), [vol], bodies, steps=10)


def small_wcsph(seed: int = 11, n_side=(12, 10, 8), spacing=0.02, tank=True, fix_density=True) -> Scene:
    sc = small_pcisph(seed, n_side, spacing, tank)
    r = spacing / 2
    h = 2 * spacing
    m = 1000.0 * spacing ** 3
    sc.settings.update(method=WCSPH, time_step=2e-4, max_vel=10.0, gas_stiffness=1000.0 * 400 / 7,
                       k_viscosity=0.02 * (h / 2) * 20 / 10 * m, surface_tension=0.001,
                       flags=FIX_WCSPH_DENSITY if fix_density else 0)
    sc.name = "small_wcsph"
    return sc


def build(sc: Scene, sim_cls=None):
    """Creates a libsph simulation for a scene in the reference's init order
    (scene.cpp:368-473: set_rect_limits, add_volume*, add_boundary*)."""
    from .simulation import FluidSimulationFactory, settings_from_dict
    phys, simset = settings_from_dict(sc.settings)
    sim = FluidSimulationFactory.build_simulation(phys, simset)
    sim.set_rect_limits(*sc.container)
    for v in sc.volumes:
        sim.add_volume(v)
    for b in sc.bodies:
        sim.add_boundary(b.samples, b.can_move, b.mass, b.R, b.com)
    return sim


def c4_batched(seed: int = 404, count: int = 512, n_min: int = 1000, n_max: int = 64000) -> List[Scene]:
    """Batched WCSPH scenes, drawn in order from one SplitMix64 stream:
    particle count N = round(exp(U(ln n_min, ln n_max))), tank sides
    U(0.5, 2) m each, Monaghan alpha U(0.005, 0.05) and c0 U(10, 40) m/s
    (mapped as in c1_wcsph_dambreak), spacing s U(0.01, 0.02) m clamped so
    that N s^3 fits in 40 % of the tank, h = 2 s, and dt = 0.25 h / c0 (the
    reference has no CFL rule; this is the stated one). The fluid is a dam
    block of shape 0.5W x 0.8H x D scaled to N lattice points, one spacing
    off the -x wall and the floor, centred in z; the tank is the integrator's
    clamp box (no boundary particles). Each scene's count is the block's
    lattice count, close to N."""
    rng = Rng(seed)
    out = []
    rho0 = 1000.0
    for k in range(count):
        u = rng.uniform(7)
        n_target = int(round(np.exp(np.log(n_min) + u[0] * (np.log(n_max) - np.log(n_min)))))
        W, H, D = (0.5 + 1.5 * u[1:4]).tolist()
        alpha = 0.005 + 0.045 * u[4]
        c0 = 10.0 + 30.0 * u[5]
        s = 0.01 + 0.01 * u[6]
        s = min(s, (0.4 * W * H * D / n_target) ** (1.0 / 3.0))
        kk = (n_target * s ** 3 / (0.4 * W * H * D)) ** (1.0 / 3.0)
        nx = max(1, int(round(kk * 0.5 * W / s)))
        ny = max(1, int(round(kk * 0.8 * H / s)))
        nz = max(1, int(round(n_target / (nx * ny))))
        nz = max(1, min(nz, int((D - 2 * s) / s)))
        size = np.array([nx, ny, nz], np.float64) * s * (1 - 1e-5)
        center = [-W / 2 + s + size[0] / 2, -H / 2 + s + size[1] / 2, 0.0]
        h, r = 2 * s, s / 2
        m = rho0 * s ** 3
        nu = alpha * (h / 2) * c0 / 10
        settings = dict(time_step=0.25 * h / c0, max_vel=10.0, fluid_particle_radius=r, fluid_support_radius=h,
                        pcisph_max_iterations=7, pcisph_error_ratio=0.01, method=WCSPH, rest_density=rho0,
                        k_viscosity=nu * m, gravity=9.81, gas_stiffness=rho0 * c0 * c0 / 7.0,
                        surface_tension=0.001, flags=FIX_WCSPH_DENSITY)
        vol = _box(size.astype(np.float32), np.array(center, np.float32), r, rng)
        assert vol.shape[0] == nx * ny * nz, (vol.shape, nx, ny, nz)
        out.append(Scene(f"batched_wcsph_{k}", settings, (W, H, D), [vol], [], steps=1000,
                         notes=f"alpha={alpha:.5f} c0={c0:.3f} target_n={n_target}"))
    return out


def build_batch(scs: List[Scene], autotune: bool = True, cache: Optional[str] = None, trials: int = 4):
    """The scenes as one libsph batch (sph_batch_*), finalised and, unless
    told otherwise, autotuned."""
    from .batch import BatchSimulation
    b = BatchSimulation()
    for sc in scs:
        parts = np.concatenate(sc.volumes) if sc.volumes else np.zeros((0, 4), np.float32)
        b.add_scene(sc.settings, sc.container, parts)
    b.finalize()
    if autotune:
        b.autotune(cache, trials)
    return b


def monaghan_k_viscosity(alpha, c0, h, mass):
    """Monaghan alpha mapped onto the reference's Laplacian viscosity (as in
    c1_wcsph_dambreak): k_viscosity = nu * m, nu = alpha (h/2) c0 / 10.
    Works on floats and on torch tensors."""
    return alpha * (h / 2) * c0 / 10 * mass


def tait_gas_stiffness(c0, rho0):
    """Tait EOS with gamma = 7: gas_stiffness = rho0 c0^2 / 7."""
    return rho0 * c0 * c0 / 7.0


@dataclass
class AdjointCase:
    scene: Scene
    v0_hidden: np.ndarray  # (n, 3) float32 initial velocities of the target run
    alpha: float
    c0: float
    steps: int
    checkpoint_every: int

    @property
    def mass(self) -> float:
        return self.scene.settings["rest_density"] * (2 * self.scene.settings["fluid_particle_radius"]) ** 3


def c5_adjoint(seed: int = 505, nx: int = 32, ny: int = 32, nz: int = 16, steps: int = 200) -> AdjointCase:
    """Differentiable WCSPH inverse problem: an nx x ny x nz block on a 0.01 m
    lattice (h = 0.02 m) centred at the origin of a 1 x 1 x 1 m clamp-box
    container, dt = 2e-4 s. Hidden alpha = 0.03, c0 = 20 m/s. The hidden
    initial velocity is, per component, a sum of 8 seeded Fourier modes
    sin(2 pi k.x / L + phase) with integer k in [-2, 2]^3, amplitude U(-1, 1)
    and L the block's largest side, scaled to max |v| = 0.5 m/s. Surface
    tension 0.001 and max_vel 10 come from config.wc; the density is the
    repaired one."""
    s, rho0, alpha, c0 = 0.01, 1000.0, 0.03, 20.0
    r, h = s / 2, 2 * s
    m = rho0 * s ** 3
    rng = Rng(seed)
    size = np.array([nx, ny, nz], np.float64) * s * (1 - 1e-5)
    vol = _box(size.astype(np.float32), np.zeros(3, np.float32), r, rng)
    settings = dict(time_step=2e-4, max_vel=10.0, fluid_particle_radius=r, fluid_support_radius=h,
                    pcisph_max_iterations=7, pcisph_error_ratio=0.01, method=WCSPH, rest_density=rho0,
                    k_viscosity=monaghan_k_viscosity(alpha, c0, h, m), gravity=9.81,
                    gas_stiffness=tait_gas_stiffness(c0, rho0), surface_tension=0.001, flags=FIX_WCSPH_DENSITY)
    L = float(max(nx, ny, nz) * s)
    x = vol[:, :3].astype(np.float64)
    v = np.zeros((vol.shape[0], 3))
    for c in range(3):
        for _ in range(8):
            k = (rng.u64() % 125)
            kv = np.array([k % 5, (k // 5) % 5, k // 25], np.float64) - 2.0
            a, ph = rng.uniform(2)
            v[:, c] += (2 * a - 1) * np.sin(2 * np.pi * (x @ kv) / L + 2 * np.pi * ph)
    v *= 0.5 / np.max(np.linalg.norm(v, axis=1))
    sc = Scene("wcsph_adjoint_16k", settings, (1.0, 1.0, 1.0), [vol], [], steps=steps)
    return AdjointCase(sc, v.astype(np.float32), alpha, c0, steps, 20)
