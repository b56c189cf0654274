"""SPH_FAST_MATH: the single-scene kernels built with the GPU's approximate
reciprocal, square root and reciprocal square root, as the reference builds
its kernels with -cl-fast-relaxed-math (pcisphsimluation.cpp:496-497,
wcsphsimluation.cpp:364). PCISPH is checked below at every size and over a
long run; WCSPH one step at a time on the small scene.

The bar, from the same state on both sides:
  * cell keys, sort permutation and neighbour lists bit-exact with the oracle
    (the neighbour search is the same code in both builds);
  * every fp32 field within 5e-5 max-relative (max |a - b| / max |b|) of the
    IEEE oracle, and the fields do differ in some bits (the fast build ran).
    `err` = max(0, rho* - rho0) (pcisph.cl:349-418) cancels the predicted
    density against the rest density and is bounded like the density it
    comes from: max |a - b| / rho0 <= 1e-5. The pressure (delta times the
    summed err), the pressure force and, from a state at rest, the new
    velocities inherit that cancellation: at the 1M scene they are 1.2e-5 to
    1.3e-5 off, the other fields below 1e-6. So the fast build does not meet
    the 1e-5 single-step bound on those fields, and the measured build is
    the IEEE one, bit-identical to the oracle;
  * over 300 steps of the 1M scene, the kinetic energy, the mean density and
    the mean density error mean(|rho - rho0|) / rho0 of a fast run stay
    within 0.5 % of the IEEE run's. (On a 960-particle dam break the runs
    part like any two rounding choices of a chaotic flow: 0.01-0.2 % after
    100 steps, up to 5 % after 300.)
"""
import numpy as np
import pytest

from fluids_amd import FAST_MATH, scenes
from helpers import build_oracle, max_rel

pytestmark = pytest.mark.gpu

TOL = 5e-5
ERR_TOL = 1e-5  # of rho0
INT_FIELDS = ["keys", "perm", "ncount", "nlist", "bncount", "bnlist"]
PCI_FIELDS = ["pos_sorted", "vel_sorted", "density", "normal", "force_other", "pressure", "err",
              "force_pressure", "pos_pred", "pos", "vel"]
# fields the approximate division and square root reach
FAST_FIELDS = ["normal", "force_other", "force_pressure"]


def _fast(sc):
    sc.settings = dict(sc.settings, flags=sc.settings.get("flags", 0) | FAST_MATH)
    return sc


def _step_and_compare(sim, orc, label, rho0):
    orc.set_state(sim.positions(), sim.velocities())
    sim.step(1)
    orc.step(1)
    assert sim.step_stats()["iterations"] == orc.iterations
    for f in INT_FIELDS:
        a, b = sim.field(f), orc.field(f)
        assert np.array_equal(a, b), f"{label}: {f} differs at {np.argwhere(a != b)[:5].ravel()}"
    report, errs = {}, {}
    for f in PCI_FIELDS:
        a, b = sim.field(f), orc.field(f)
        if a.ndim == 2:
            a, b = a[:, :3], b[:, :3]
        errs[f] = max_rel(a, b) if f != "err" else float(np.abs(a.astype(np.float64) - b).max() / rho0)
        report[f] = (f"{errs[f]:.2e}", int(np.sum(a.view(np.uint32) != b.view(np.uint32))))
    print(label, "max-rel, differing words:", report, flush=True)
    for f, e in errs.items():
        tol = ERR_TOL if f == "err" else TOL
        assert e <= tol, f"{label}: {f} max-rel {e:.3e} > {tol}"
    return report


@pytest.mark.parametrize("cube", [False, True])
def test_fast_math_single_step_small(cube):
    sc = _fast(scenes.small_pcisph(cube=cube))
    print("case fast-math pcisph small", sc.n_fluid, sc.n_boundary, "cube", cube, flush=True)
    sim = scenes.build(sc)
    orc = build_oracle(scenes.small_pcisph(cube=cube))
    differing = 0
    for step in range(3):
        rep = _step_and_compare(sim, orc, f"small step {step}", sc.settings["rest_density"])
        differing += sum(rep[f][1] for f in FAST_FIELDS)
    assert differing > 0, "no field differs from the IEEE oracle: the fast build did not run"


def test_fast_math_single_step_1m():
    sc = _fast(scenes.c3_pcisph_throughput(1_000_000))
    print("case fast-math pcisph 1M", sc.n_fluid, sc.n_boundary, flush=True)
    sim = scenes.build(sc)
    orc = build_oracle(scenes.c3_pcisph_throughput(1_000_000))
    rep = _step_and_compare(sim, orc, "1M step", sc.settings["rest_density"])
    assert sum(rep[f][1] for f in FAST_FIELDS) > 0


WC_FIELDS = ["pos_sorted", "vel_sorted", "vel_half_sorted", "density", "pressure", "normal", "accel", "pos",
             "vel", "vel_half"]


def test_fast_math_wcsph_single_step():
    sc = _fast(scenes.small_wcsph())
    print("case fast-math wcsph small", sc.n_fluid, sc.n_boundary, flush=True)
    sim = scenes.build(sc)
    orc = build_oracle(scenes.small_wcsph())
    differing = 0
    for step in range(3):
        pos, vel, vh = sim.positions(), sim.velocities(), sim.half_velocities()
        orc.set_state(pos, vel, vh)
        sim.set_positions(pos)
        sim.set_velocities(vel)
        sim.set_half_velocities(vh)
        sim.step(1)
        orc.step(1)
        for f in INT_FIELDS:
            a, b = sim.field(f), orc.field(f)
            assert np.array_equal(a, b), f"wcsph step {step}: {f} differs"
        report = {}
        for f in WC_FIELDS:
            a, b = sim.field(f), orc.field(f)
            if a.ndim == 2:
                a, b = a[:, :3], b[:, :3]
            e = max_rel(a, b)
            report[f] = (f"{e:.2e}", int(np.sum(a.view(np.uint32) != b.view(np.uint32))))
            assert e <= TOL, f"wcsph step {step}: {f} max-rel {e:.3e} > {TOL}"
        print(f"wcsph step {step} max-rel, differing words:", report, flush=True)
        differing += report["normal"][1] + report["accel"][1]
    assert differing > 0, "no field differs from the IEEE oracle: the fast build did not run"


def _kinetic_and_density(sim, mass, rho0):
    v = sim.velocities()[:, :3].astype(np.float64)
    ke = 0.5 * mass * float((v * v).sum())
    rho = sim.field("density").astype(np.float64)
    return ke, float(rho.mean()), float(np.abs(rho - rho0).mean() / rho0)


def test_fast_math_long_run_aggregates():
    """The 1M scene's boxes collapsing onto the tank floor, 300 steps of each
    build from the same state: kinetic energy, mean density and mean density
    error agree within 0.5 % every 100 steps."""
    ref_sc = scenes.c3_pcisph_throughput(1_000_000)
    print("case fast-math long run", ref_sc.n_fluid, ref_sc.n_boundary, flush=True)
    ieee = scenes.build(ref_sc)
    fast = scenes.build(_fast(scenes.c3_pcisph_throughput(1_000_000)))
    fast.set_positions(ieee.positions())
    fast.set_velocities(ieee.velocities())
    s = ref_sc.settings
    mass = s["rest_density"] * (2 * s["fluid_particle_radius"]) ** 3 / 1.15  # pcisphsimluation.cpp:236
    rho0 = s["rest_density"]
    worst = []
    for block in range(3):
        ieee.step(100)
        fast.step(100)
        a = _kinetic_and_density(ieee, mass, rho0)
        b = _kinetic_and_density(fast, mass, rho0)
        assert min(a) > 0
        rel = [abs(x - y) / abs(x) for x, y in zip(a, b)]
        print(f"step {100 * (block + 1)}: ieee {a} fast {b} rel {['%.2e' % r for r in rel]}", flush=True)
        worst.append(max(rel))
    assert max(worst) <= 5e-3, worst
    assert not np.array_equal(ieee.positions(), fast.positions())
