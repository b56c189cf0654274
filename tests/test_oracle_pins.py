"""Closed forms and invariants for the oracle's force, pressure and integrator
functions (CPU only).

Each case is a configuration small or symmetric enough that the reference's
equations reduce to a closed expression written out here in float64 from the
geometry alone: a line of three particles, a pair at a chosen distance (one
case per branch of the surface-tension spline), a lone particle, a lone
particle over a planar boundary lattice, or an invariant (total momentum,
translation invariance). Field values are compared with those expressions;
the kernel constants they use are pinned to their closed forms in
test_oracle.py and again below for the WCSPH set.

Reference equations (kernels/...):
  normals.cl:23-70            n_i = G m h sum_j r_ij (h^2 - r^2)^2 / rho_j
  pcisph.cl:134-222 and       viscosity, curvature and cohesion (Akinci)
  surface_tension.h:20-80
  pcisph.cl:349-418           rho*_i, err = max(0, rho* - rho0), p += delta err
  pcisph.cl:420-504           pressure force, fluid and boundary terms
  boundary.cl:61-118          per-sample fluid force on a body
  wcsph.cl:32-79, 81-173      WCSPH density, Tait pressure, acceleration
  wcsph.cl:176-224            leap-frog with the Q12 position update
"""
import math

import numpy as np
import pytest

import oracle

f32 = np.float32

PCI = dict(time_step=1e-3, max_vel=20.0, fluid_particle_radius=0.01, fluid_support_radius=0.04,
           pcisph_max_iterations=7, pcisph_error_ratio=0.01, method=1, rest_density=1000.0,
           k_viscosity=1.7, gravity=9.8, gas_stiffness=500.0, surface_tension=1.0, flags=0)
WC = dict(time_step=2e-4, max_vel=20.0, fluid_particle_radius=0.0175, fluid_support_radius=0.04,
          pcisph_max_iterations=7, pcisph_error_ratio=0.01, method=0, rest_density=1000.0,
          k_viscosity=0.3, gravity=9.81, gas_stiffness=1000.0, surface_tension=0.8, flags=1)


def f4(p):
    p = np.asarray(p, np.float32).reshape(-1, 3)
    return np.concatenate([p, np.ones((len(p), 1), np.float32)], 1)


def run(settings, pos, vel=None, vh=None, walls=(), container=(2.0, 1.0, 1.0), steps=1):
    """walls: (samples, mass) pairs, added as bodies with the identity transform
    (mass > 0 makes the body receive a fluid wrench)."""
    o = oracle.Oracle(settings)
    o.set_rect_limits(*container)
    for samples, mass in walls:
        o.add_boundary(f4(samples), mass > 0, mass, np.eye(3, dtype=np.float32), np.zeros(3, np.float32))
    o.set_state(f4(pos), None if vel is None else f4(vel), None if vh is None else f4(vh))
    o.step(steps)
    return o


def by_particle(o, name):
    """A sorted-order field re-indexed by input particle."""
    a = o.field(name)
    out = np.empty_like(a)
    out[o.field("perm")] = a
    return out


def spline(r, h, c2):
    """st_kernel (surface_tension.h:23-36)."""
    if r <= 0.5 * h:
        return 2.0 * (h - r) ** 3 * r ** 3 - c2
    if r <= h:
        return (h - r) ** 3 * r ** 3
    return 0.0


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


# A unit direction with three non-zero components, so every component of each
# vector term is exercised: (2, 1, -2) / 3.
U = np.array([2.0, 1.0, -2.0]) / 3.0
BASE = np.array([0.11, 0.07, -0.05])


def pci_consts(settings):
    return oracle.Oracle(settings).constants()


# ---------------------------------------------------------------- normals
def test_normals_three_particles_in_a_line():
    """i, j, k on a line with |ij| = a < h, |jk| = b < h, |ik| = a + b > h.
    Densities (self included): rho_i = W M (h^6 + (h^2 - a^2)^3),
    rho_j = W M (h^6 + (h^2 - a^2)^3 + (h^2 - b^2)^3), rho_k likewise with b.
    Then n_i = G M h (x_i - x_j) (h^2 - a^2)^2 / rho_j, etc. A density taken
    from the wrong particle, a missing factor or a wrong power fails."""
    c = pci_consts(PCI)
    H, M, W, G = (float(c[k]) for k in ("pci_h_d", "pci_mass_d", "pci_default_eval", "pci_default_grad"))
    p = np.array([BASE, BASE + 0.015 * U, BASE + 0.045 * U], np.float32)
    o = run(dict(PCI, gravity=0.0), p)
    x = p.astype(np.float64)
    r = lambda i, j: x[i] - x[j]
    q = lambda i, j: H * H - r(i, j) @ r(i, j)
    assert q(0, 2) < 0 < q(0, 1) and q(1, 2) > 0
    rho = W * M * np.array([H ** 6 + q(0, 1) ** 3, H ** 6 + q(0, 1) ** 3 + q(1, 2) ** 3, H ** 6 + q(1, 2) ** 3])
    assert rel(by_particle(o, "density"), rho) < 2e-6
    n = G * M * H * np.array([
        r(0, 1) * q(0, 1) ** 2 / rho[1],
        r(1, 0) * q(1, 0) ** 2 / rho[0] + r(1, 2) * q(1, 2) ** 2 / rho[2],
        r(2, 1) * q(2, 1) ** 2 / rho[1],
    ])
    got = by_particle(o, "normal")[:, :3]
    for i in range(3):
        assert rel(got[i], n[i]) < 1e-5, (i, got[i], n[i])
    # the end particles' normals point along the line into the group
    assert got[0] @ U > 0 and got[2] @ U < 0


def lattice(n, s, origin):
    g = [np.arange(k) * s for k in n]
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + np.asarray(origin)


def test_normals_lattice_interior_zero_and_free_surface_inward():
    """Unperturbed 9 x 5 x 9 lattice at spacing 2r: the centre particle's
    neighbourhood is point-symmetric, so its normal vanishes (to rounding);
    the centre of the top face has tangential components that vanish by the
    mirror symmetries and a normal component pointing into the fluid (-y),
    since grad W_poly6 < 0 along r_ij = x_i - x_j (normals.cl:66)."""
    s = 0.02
    p = lattice((9, 5, 9), s, (-0.08 + 0.003, -0.2 + 0.001, -0.08 + 0.002)).astype(np.float32)
    o = run(dict(PCI, gravity=0.0), p)
    n = by_particle(o, "normal")[:, :3].astype(np.float64)
    centre = 4 * 45 + 2 * 9 + 4
    top = 4 * 45 + 4 * 9 + 4
    assert np.allclose(p[top] - p[centre], [0, 2 * s, 0], atol=1e-6)
    scale = np.abs(n[top]).max()
    assert np.abs(n[centre]).max() < 1e-5 * scale
    assert abs(n[top][0]) < 1e-5 * scale and abs(n[top][2]) < 1e-5 * scale
    assert n[top][1] < 0


# ---------------------------------------------------- viscosity and surface tension
@pytest.mark.parametrize("dist", [0.015, 0.03])  # r <= h/2 and h/2 < r < h branches of st_kernel
def test_other_force_pair_closed_form(dist):
    """Two particles at distance r with velocities v_i, v_j. Both have
    rho = W M (h^6 + (h^2 - r^2)^3) and n_j = -n_i = -G M h r_ij (h^2-r^2)^2 / rho.
    compute_initial_forces (pcisph.cl:134-222) then gives
    F_i = k m^2 Wl / rho * (v_j - v_i)(h - r) / rho
          - sigma m [ 2 n_i K + m c1 K r^_ij st(r) ],  K = 2 rho0 / (2 rho).
    The self pair adds nothing (v_i - v_i = 0, n_i - n_i = 0, normalize(0) = 0)."""
    c = pci_consts(PCI)
    H, M, W, G = (float(c[k]) for k in ("pci_h_d", "pci_mass_d", "pci_default_eval", "pci_default_grad"))
    VL, C1, C2, R0 = (float(c[k]) for k in ("pci_visc_lapl", "pci_st_main", "pci_st_term", "pci_rho0_d"))
    p = np.array([BASE, BASE + dist * U], np.float32)
    v = np.array([[0.3, -0.2, 0.1], [-0.1, 0.25, 0.4]], np.float32)
    o = run(PCI, p, v)
    x = p.astype(np.float64)
    rv = x[0] - x[1]
    r = math.sqrt(rv @ rv)
    rho = W * M * (H ** 6 + (H * H - r * r) ** 3)
    ni = G * M * H * rv * (H * H - r * r) ** 2 / rho
    K = 2 * R0 / (2 * rho)
    k, sigma = PCI["k_viscosity"], PCI["surface_tension"]
    st = spline(r, H, C2)
    vd = v.astype(np.float64)
    Fi = (k * M * M * VL / rho * (vd[1] - vd[0]) * (H - r) / rho
          - sigma * M * (2 * ni * K + M * C1 * K * (rv / r) * st))
    Fj = (k * M * M * VL / rho * (vd[0] - vd[1]) * (H - r) / rho
          - sigma * M * (-2 * ni * K + M * C1 * K * (-rv / r) * st))
    got = by_particle(o, "force_other")[:, :3]
    assert rel(got[0], Fi) < 1e-5, (got[0], Fi)
    assert rel(got[1], Fj) < 1e-5, (got[1], Fj)
    # cohesion alone: attractive (towards j) below h/2 only where st > 0
    o2 = run(dict(PCI, k_viscosity=0.0), p, v)
    coh = by_particle(o2, "force_other")[0, :3].astype(np.float64) + sigma * M * 2 * ni * K
    want = -sigma * M * M * C1 * K * (rv / r) * st
    assert rel(coh, want) < 1e-4


def test_other_force_conserves_momentum():
    """Viscosity (v_j - v_i)/(rho_i rho_j), curvature (n_i - n_j) K_ij and
    cohesion r^_ij st K_ij are each antisymmetric in i, j; with symmetric
    lists (below the cap, no walls) the forces sum to zero."""
    rng = np.random.default_rng(5)
    p = lattice((7, 6, 5), 0.02, (-0.06, -0.05, -0.04))
    p = (p + rng.uniform(-0.004, 0.004, p.shape)).astype(np.float32)
    v = rng.uniform(-0.5, 0.5, p.shape).astype(np.float32)
    o = run(PCI, p, v)
    assert o.field("ncount").max() < 50
    f = o.field("force_other")[:, :3].astype(np.float64)
    assert np.abs(f).max() > 0
    assert np.abs(f.sum(0)).max() < 2e-6 * np.abs(f).sum()


# ------------------------------------------------------------ update_pressure
LONE = dict(PCI, fluid_particle_radius=0.02, gravity=0.0, surface_tension=0.0, pcisph_error_ratio=-1.0)


@pytest.mark.parametrize("iters", [1, 3, 7])
def test_update_pressure_lone_particles(iters):
    """Particles farther apart than h: each one's list is itself, so with
    g = 0, v = 0 and sigma = 0 nothing moves, rho* = rho = W M h^6 in every
    correction, err = rho - rho0 (> 0 at h = 2r) and, after k corrections,
    p = k delta err (pcisph.cl:349-418). The pressure force skips r < 1e-5,
    so it stays exactly zero."""
    s = dict(LONE, pcisph_max_iterations=iters)
    c = pci_consts(s)
    H, M, W, R0, delta = (float(c[k]) for k in ("pci_h_d", "pci_mass_d", "pci_default_eval", "pci_rho0_d",
                                                 "pci_delta"))
    p = np.array([[0.1, 0.0, 0.0], [0.2, 0.05, 0.0], [-0.3, 0.1, 0.2]], np.float32)
    o = run(s, p)
    assert o.iterations == iters
    rho = W * M * H ** 6
    assert rho > R0
    assert rel(o.field("density"), np.full(3, rho)) < 1e-6
    assert rel(o.field("err"), np.full(3, rho - R0)) < 1e-5
    assert rel(o.field("pressure"), np.full(3, iters * delta * (rho - R0))) < 1e-5
    assert (o.field("force_pressure") == 0).all()


def test_update_pressure_is_translation_invariant():
    """Uniform velocity and gravity, sigma = 0, one correction: F_o = 0 and
    F_p = 0 in the prediction, so every predicted position is the stored one
    shifted by the same vector; rho* of the predicted state equals the
    initial density (to rounding), err = max(0, rho - rho0), p = delta err.
    A neighbour read at its current instead of its predicted position fails."""
    s = dict(PCI, surface_tension=0.0, pcisph_max_iterations=1)
    rng = np.random.default_rng(9)
    p = lattice((8, 7, 6), 0.019, (-0.07, -0.06, -0.05))
    p = (p + rng.uniform(-0.002, 0.002, p.shape)).astype(np.float32)
    v = np.tile(np.array([[0.7, -0.3, 0.5]], np.float32), (len(p), 1))
    o = run(s, p, v)
    c = o.constants()
    rho, err = o.field("density").astype(np.float64), o.field("err").astype(np.float64)
    R0 = float(c["pci_rho0_d"])
    assert (rho > R0).sum() > 50 and (rho < R0).sum() > 50
    assert np.abs(err - np.maximum(0, rho - R0)).max() < 2e-5 * R0
    shift = o.field("pos_pred")[:, :3].astype(np.float64) - o.field("pos_sorted")[:, :3]
    assert np.abs(shift - shift[0]).max() < 1e-6 and np.abs(shift[0]).max() > 1e-4
    assert np.array_equal(o.field("pressure"), (o.field("err") * f32(c["pci_delta"])).astype(np.float32))


# ------------------------------------------------------------- boundary terms
S_WALL = 0.017  # no lattice offset lies at distance h = 0.04 from a site or from the test particle
D_WALL = 0.015


def wall_lattice(y0=-0.2, half=10):
    k = np.arange(-half, half + 1) * S_WALL
    X, Z = np.meshgrid(k, k, indexing="ij")
    return np.stack([X.ravel(), np.full(X.size, y0), Z.ravel()], 1).astype(np.float32)


def planar_psi(h, w_eval, rho0):
    """psi of an interior site of the square lattice (boundary.cl:4-31)."""
    k = np.array([(a, b) for a in range(-4, 5) for b in range(-4, 5)], np.float64)
    r2 = (k ** 2).sum(1) * S_WALL * S_WALL
    return rho0 / (w_eval * ((h * h - r2[r2 < h * h]) ** 3).sum())


def over_wall(settings, mass=1.0):
    w = wall_lattice()
    p = np.array([[0.0, -0.2 + D_WALL, 0.0]], np.float32)
    o = run(settings, p, walls=[(w, mass)])
    x = p[0].astype(np.float64)
    d = x - w.astype(np.float64)
    r = np.sqrt((d ** 2).sum(1))
    near = np.flatnonzero(r < settings["fluid_support_radius"])
    return o, w, x, d[near], r[near], near


def test_boundary_pressure_force_lone_particle_closed_form():
    """One fluid particle at height d above a site of a planar boundary
    lattice, g = v = sigma = 0, one correction. With psi from the planar
    closed form: rho = W (M h^6 + sum_b psi (h^2 - r_b^2)^3),
    p = delta max(0, rho - rho0), C = p / rho^2 and
    F_p = -M Wsp sum_b psi 2C r_b / |r_b| (h - |r_b|)^2 (pcisph.cl:486-500):
    normal to the wall, away from it, tangential parts cancel."""
    s = dict(LONE, pcisph_max_iterations=1)
    o, w, x, d, r, near = over_wall(s)
    c = o.constants()
    H, M, W, R0 = (float(c[k]) for k in ("pci_h_d", "pci_mass_d", "pci_default_eval", "pci_rho0_d"))
    WPG, delta = float(c["pci_pressure_grad"]), float(c["pci_delta"])
    psi = planar_psi(float(c["b_h_d"]), float(c["b_poly6_eval"]), 1000.0)
    assert rel(o.field("bphi")[near], np.full(len(near), psi)) < 1e-5
    assert o.field("bncount")[0] == len(near) > 4
    rho = W * (M * H ** 6 + (psi * (H * H - r * r) ** 3).sum())
    assert rel(o.field("density"), [rho]) < 1e-6
    pr = delta * max(0.0, rho - R0)
    assert rel(o.field("pressure"), [pr]) < 1e-5
    C = pr / rho ** 2
    F = -M * WPG * (psi * 2 * C * (H - r) ** 2 / r)[:, None] * d
    F = F.sum(0)
    got = o.field("force_pressure")[0, :3].astype(np.float64)
    assert F[1] > 0 and abs(F[0]) < 1e-9 * F[1] and abs(F[2]) < 1e-9 * F[1]
    assert abs(got[1] - F[1]) < 1e-5 * F[1]
    assert abs(got[0]) < 1e-6 * F[1] and abs(got[2]) < 1e-6 * F[1]


def test_boundary_fluid_force_closed_form_and_third_law():
    """compute_fluid_force (boundary.cl:61-118) on the same configuration with
    the wall as a body of mass > 0: each site within h gets
    f_b = m psi_b 2 C (x_i - x_b) / |r| Wsp (h - |r|)^2 and every other site 0;
    summed over the wall, -(m / M) times the particle's boundary pressure
    force (the reference's m is the full fp32 mass, M the 6-decimal literal)."""
    s = dict(LONE, pcisph_max_iterations=1)
    o, w, x, d, r, near = over_wall(s, mass=2.5)
    c = o.constants()
    SG, HB, M = float(c["b_spiky_grad"]), float(c["b_h_d"]), float(c["pci_mass_d"])
    m = float(c["pci_mass"])
    psi = o.field("bphi").astype(np.float64)
    rho = float(o.field("density")[0])
    C = float(o.field("pressure")[0]) / rho ** 2
    fb = np.zeros((len(w), 3))
    fb[near] = m * (psi[near] * 2 * C * SG * (HB - r) ** 2 / r)[:, None] * d
    got = o.field("bforce")[:, :3].astype(np.float64)
    assert np.count_nonzero(np.abs(got).sum(1)) == len(near)
    assert np.abs(got - fb).max() < 1e-5 * np.abs(fb).max()
    fp = o.field("force_pressure")[0, :3].astype(np.float64)
    total = got.sum(0)
    assert total[1] < 0
    assert np.abs(total + (m / M) * fp).max() < 1e-5 * np.abs(fp).max()
    F, T, F64, T64 = o.body_wrench(0)
    assert np.abs(F - total).max() < 1e-6 * np.abs(total).max()


def test_boundary_third_law_for_a_fluid_group():
    """A few fluid particles over the wall (fluid-fluid pressure terms cancel
    in pairs): sum_b f_b = -(m / M) sum_i F_p,i."""
    s = dict(PCI, gravity=0.0, surface_tension=0.0, pcisph_max_iterations=3)
    w = wall_lattice()
    rng = np.random.default_rng(2)
    p = lattice((4, 2, 4), 0.02, (-0.03, -0.2 + 0.012, -0.03))
    p = (p + rng.uniform(-0.002, 0.002, p.shape)).astype(np.float32)
    o = run(s, p, walls=[(w, 1.0)])
    c = o.constants()
    assert o.field("ncount").max() < 50 and o.field("bncount").max() < 50
    fp = o.field("force_pressure")[:, :3].astype(np.float64).sum(0)
    fb = o.field("bforce")[:, :3].astype(np.float64).sum(0)
    ratio = float(c["pci_mass"]) / float(c["pci_mass_d"])
    assert np.abs(fp).max() > 0
    assert np.abs(fb + ratio * fp).max() < 1e-5 * np.abs(o.field("force_pressure")).sum()


# ------------------------------------------------------------------- WCSPH
def wc_consts_closed(s):
    h = float(f32(s["fluid_support_radius"]))
    m = float(f32(s["rest_density"] * f32(2 * f32(s["fluid_particle_radius"])) ** 3))
    return dict(
        h=h, m=m,
        poly6=315 / (64 * math.pi * h ** 9), grad=-945 / (32 * math.pi * h ** 9),
        spiky=-45 / (math.pi * h ** 6), visc=45 / (math.pi * h ** 6),
        st_main=32 / (math.pi * h ** 9), st_term=h ** 6 / 64,
        m_d=float(f32(float(f"{f32(m):.6f}"))), h_d=float(f32(float(f"{f32(h):.6f}"))),
    )


def test_wcsph_constants_closed_form():
    c = oracle.Oracle(WC).constants()
    k = wc_consts_closed(WC)
    for a, b in [("wc_mass", "m"), ("wc_poly6_eval", "poly6"), ("wc_poly6_grad", "grad"),
                 ("wc_spiky_grad", "spiky"), ("wc_visc_lapl", "visc"), ("wc_st_main", "st_main"),
                 ("wc_st_term", "st_term"), ("wc_mass_d", "m_d"), ("wc_h_d", "h_d")]:
        assert abs(c[a] - k[b]) <= 2e-6 * abs(k[b]), (a, c[a], k[b])


@pytest.mark.parametrize("dist", [0.015, 0.03])
def test_wcsph_acceleration_pair_closed_form(dist):
    """Two particles at distance r (compute_acceleration, wcsph.cl:81-173, with
    the r^2 density of SPH_FIX_WCSPH_DENSITY so that rho > rho0):
    rho = W315 m (h^6 + (h^2 - r^2)^3), p = k ((rho/rho0)^7 - 1), C = p / rho^2,
    n_i = G M_d h_d r_ij (h_d^2 - r^2)^2 / rho,
    a_i = g - sigma [2 n_i K + c1 K r^ st(r)] - m Wsp (h - r)^2 2C r^
          + Wl k (v_j - v_i)(h - r) / rho,   K = 2 rho0 / (2 rho)."""
    k = wc_consts_closed(WC)
    h, m = k["h"], k["m"]
    p = np.array([BASE, BASE + dist * U], np.float32)
    v = np.array([[0.3, -0.2, 0.1], [-0.1, 0.25, 0.4]], np.float32)
    o = run(WC, p, v, np.zeros_like(v))
    x = p.astype(np.float64)
    rv = x[0] - x[1]
    r = math.sqrt(rv @ rv)
    rho = k["poly6"] * m * (h ** 6 + (h * h - r * r) ** 3)
    r0 = WC["rest_density"]
    pr = max(0.0, WC["gas_stiffness"] * ((rho / r0) ** 7 - 1))
    assert pr > 0
    assert rel(by_particle(o, "density"), [rho, rho]) < 2e-6
    assert rel(by_particle(o, "pressure"), [pr, pr]) < 2e-5
    C = pr / rho ** 2
    Hd, Md = k["h_d"], k["m_d"]
    ni = k["grad"] * Md * Hd * rv * (Hd * Hd - r * r) ** 2 / rho
    assert rel(by_particle(o, "normal")[0, :3], ni) < 1e-5
    K = 2 * r0 / (2 * rho)
    st = spline(r, h, k["st_term"])
    g = np.array([0.0, -WC["gravity"], 0.0])
    vd = v.astype(np.float64)
    sig, kv = WC["surface_tension"], WC["k_viscosity"]

    def acc(sgn, vi, vj):
        u = sgn * rv / r
        return (g - sig * (2 * sgn * ni * K + k["st_main"] * K * u * st)
                - m * k["spiky"] * (h - r) ** 2 * 2 * C * u
                + k["visc"] * kv * (vj - vi) * (h - r) / rho)
    got = by_particle(o, "accel")[:, :3]
    a0, a1 = acc(1, vd[0], vd[1]), acc(-1, vd[1], vd[0])
    assert rel(got[0], a0) < 1e-5, (got[0], a0)
    assert rel(got[1], a1) < 1e-5, (got[1], a1)
    # the pressure part alone pushes the pair apart
    o2 = run(dict(WC, gravity=0.0, k_viscosity=0.0, surface_tension=0.0), p, v, np.zeros_like(v))
    a = by_particle(o2, "accel")[:, :3].astype(np.float64)
    assert a[0] @ rv > 0 and a[1] @ rv < 0


def test_wcsph_three_particles_in_a_line():
    """i, j, k on a line as in the PCISPH normals case, now with WCSPH and
    distinct densities rho_i != rho_j != rho_k: each normal divides by the
    neighbour's density, and the end particles' accelerations are
    a_i = g - sigma [(n_i - n_j) K_ij + c1 K_ij r^ st(r)]
          - m Wsp (h - r)^2 (C_i + C_j) r^ + Wl k (v_j - v_i)(h - r) / rho_j,
    K_ij = 2 rho0 / (rho_i + rho_j) (wcsph.cl:81-173)."""
    k = wc_consts_closed(WC)
    h, m, Hd, Md = k["h"], k["m"], k["h_d"], k["m_d"]
    p = np.array([BASE, BASE + 0.015 * U, BASE + 0.045 * U], np.float32)
    v = np.array([[0.3, -0.2, 0.1], [-0.1, 0.25, 0.4], [0.05, 0.0, -0.3]], np.float32)
    o = run(WC, p, v, np.zeros_like(v))
    x, vd = p.astype(np.float64), v.astype(np.float64)
    rv = lambda i, j: x[i] - x[j]
    dist = lambda i, j: math.sqrt(rv(i, j) @ rv(i, j))
    pairs = {0: [1], 1: [0, 2], 2: [1]}
    assert dist(0, 2) > h
    rho = np.array([k["poly6"] * m * (h ** 6 + sum((h * h - dist(i, j) ** 2) ** 3 for j in pairs[i]))
                    for i in range(3)])
    assert len(set(np.round(rho, 6))) == 3
    assert rel(by_particle(o, "density"), rho) < 2e-6
    r0 = WC["rest_density"]
    pr = WC["gas_stiffness"] * ((rho / r0) ** 7 - 1)
    assert (pr > 0).all()
    n = np.array([sum(k["grad"] * Md * Hd * rv(i, j) * (Hd * Hd - dist(i, j) ** 2) ** 2 / rho[j]
                      for j in pairs[i]) for i in range(3)])
    got_n = by_particle(o, "normal")[:, :3]
    for i in range(3):
        assert rel(got_n[i], n[i]) < 1e-5, i
    C = pr / rho ** 2
    g = np.array([0.0, -WC["gravity"], 0.0])
    sig, kv = WC["surface_tension"], WC["k_viscosity"]
    got = by_particle(o, "accel")[:, :3]
    for i in range(3):
        a = g.copy()
        for j in pairs[i]:
            r = dist(i, j)
            u = rv(i, j) / r
            K = 2 * r0 / (rho[i] + rho[j])
            a += -sig * ((n[i] - n[j]) * K + k["st_main"] * K * u * spline(r, h, k["st_term"]))
            a += -m * k["spiky"] * (h - r) ** 2 * (C[i] + C[j]) * u
            a += k["visc"] * kv * (vd[j] - vd[i]) * (h - r) / rho[j]
        assert rel(got[i], a) < 1e-5, (i, got[i], a)


def test_wcsph_pressure_acceleration_conserves_momentum():
    """Without walls, viscosity, tension and gravity the WCSPH acceleration is
    the symmetric pressure term -m sum_j (C_i + C_j) grad W_ij: equal masses,
    so the accelerations sum to zero."""
    s = dict(WC, gravity=0.0, k_viscosity=0.0, surface_tension=0.0)
    rng = np.random.default_rng(4)
    p = lattice((7, 6, 6), 0.02, (-0.06, -0.05, -0.05))
    p = (p + rng.uniform(-0.003, 0.003, p.shape)).astype(np.float32)
    o = run(s, p, np.zeros_like(p), np.zeros_like(p))
    assert (o.field("pressure") > 0).sum() > 100 and o.field("ncount").max() < 50
    a = o.field("accel")[:, :3].astype(np.float64)
    assert np.abs(a.sum(0)).max() < 2e-6 * np.abs(a).sum()


def test_wcsph_boundary_acceleration_lone_particle():
    """One WCSPH particle over the planar lattice: rho = W315 (m h^6 +
    sum_b psi (h^2 - r^2)^3), and a = -sum_b psi r_b (Wsp / r)(h - r)^2 2C
    (wcsph.cl:130-138): away from the wall; the wall's summed fluid force is
    -m a (boundary.cl:61-118)."""
    s = dict(WC, gravity=0.0, surface_tension=0.0, k_viscosity=0.0)
    o, w, x, d, r, near = over_wall(s, mass=1.0)
    k = wc_consts_closed(s)
    c = o.constants()
    psi = planar_psi(float(c["b_h_d"]), float(c["b_poly6_eval"]), 1000.0)
    h, m = k["h"], k["m"]
    rho = k["poly6"] * (m * h ** 6 + (psi * (h * h - r * r) ** 3).sum())
    assert rel(o.field("density"), [rho]) < 2e-6
    pr = s["gas_stiffness"] * ((rho / s["rest_density"]) ** 7 - 1)
    C = pr / rho ** 2
    a = -(psi * k["spiky"] * (h - r) ** 2 / r * 2 * C)[:, None] * d
    a = a.sum(0)
    got = o.field("accel")[0, :3].astype(np.float64)
    assert a[1] > 0
    assert abs(got[1] - a[1]) < 2e-5 * a[1] and abs(got[0]) < 1e-6 * a[1] and abs(got[2]) < 1e-6 * a[1]
    total = o.field("bforce")[:, :3].astype(np.float64).sum(0)
    assert np.abs(total + m * got).max() < 2e-5 * np.abs(m * got).max()


# ---------------------------------------------------------- WCSPH leap-frog (Q12)
LF = dict(WC, time_step=0.01, max_vel=3.0, surface_tension=0.0)


def leapfrog(x, vh, dt, g=-9.81, maxv=3.0, container=(2.0, 1.0, 1.0)):
    """Reference closed form of one step for a lone particle (a = g):
    vh' = vh + dt a; x' = x + dt clamp(vh' + dt a) (the Q12 position update,
    wcsph.cl:206-207, advances by vh + 2 dt a, not the leap-frog's vh + dt a);
    v = vh' + dt a / 2; both velocities clamped to +-max_vel on output."""
    a = np.array([0.0, g, 0.0])
    vh1 = vh + dt * a
    xn = x + dt * np.clip(vh1 + dt * a, -maxv, maxv)
    return xn, np.clip(vh1, -maxv, maxv), np.clip(vh1 + dt * a * 0.5, -maxv, maxv)


@pytest.mark.parametrize("vh0", [(0.5, 0.2, -0.1), (5.0, 0.0, -4.0)])
def test_wcsph_leapfrog_lone_particle(vh0):
    x0 = np.array([[0.1, 0.2, -0.1]], np.float32)
    vh = np.array([vh0], np.float32)
    o = run(LF, x0, np.zeros_like(vh), vh)
    xn, vh1, vt = leapfrog(x0[0].astype(np.float64), vh[0].astype(np.float64), 0.01)
    assert np.abs(o.field("pos")[0, :3] - xn).max() < 2e-7
    assert np.abs(o.field("vel_half")[0, :3] - vh1).max() < 2e-6
    assert np.abs(o.field("vel")[0, :3] - vt).max() < 2e-6
    # the textbook leap-frog position differs by dt^2 g = 9.8e-4 in y
    if vh0[0] < 3:
        assert abs(o.field("pos")[0, 1] - (0.2 + 0.01 * (0.2 - 0.0981))) > 5e-4


def test_wcsph_leapfrog_floor_reflection():
    """Crossing the floor y = -h/2 within the step: the position is clamped to
    it and vh' is reflected with restitution d / (dt |vh'|) (wcsph.cl:209-219);
    v = vh'' + dt a / 2."""
    x0 = np.array([[0.0, -0.499, 0.0]], np.float32)
    vh = np.array([[0.0, -1.0, 0.0]], np.float32)
    o = run(LF, x0, np.zeros_like(vh), vh)
    dt, g = 0.01, -9.81
    vh1 = float(vh[0, 1]) + dt * g
    y = float(x0[0, 1]) + dt * (vh1 + dt * g)
    d = -0.5 - y
    assert d > 0
    rest = d / (dt * abs(vh1))
    vh2 = vh1 - (1 + rest) * vh1
    assert o.field("pos")[0, 1] == f32(-0.5)
    assert abs(o.field("vel_half")[0, 1] - vh2) < 1e-5
    assert abs(o.field("vel")[0, 1] - (vh2 + dt * g * 0.5)) < 1e-5
