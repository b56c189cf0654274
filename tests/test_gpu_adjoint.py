"""The hand-written WCSPH adjoint (sph_adjoint_*, fluids_amd.adjoint).

* Its forward pass is the solver's trajectory, bit for bit.
* On small scenes its gradients (initial positions, velocities, half-step
  velocities, k_viscosity, gas_stiffness, surface_tension) match torch
  autograd through a float64 restatement of the same step (torch_wcsph.py),
  one scene in the bulk and one whose particles hit a wall of the clamp box.
* On the 16 384-particle inverse problem (200 steps, checkpoints every 20),
  the gradient of the mean squared distance to the target positions with
  respect to the initial velocities, alpha and c0 (through
  torch.autograd.Function) matches float64 central finite differences of the
  oracle's double-precision rollout on 64 seeded directions.

The directions are smooth random velocity fields (4 Fourier modes per
component, like the hidden field). Along white-noise directions central
differences do not converge: the loss is only piecewise smooth at particle
scale (the viscosity kernel h - q has a kink at the cutoff, the pressure a
kink at rest density), and the finite difference of one such direction moves
by 2x between steps of 1e-3 and 1e-5 while the adjoint stays inside that
spread.

Each direction's error |adjoint - fd| is bounded by 1 % of the RMS of the
directional derivatives (a direction nearly orthogonal to the gradient has a
tiny derivative, so its relative error measures only the noise floor), and
the median relative error by 0.5 %.
"""
import os

import numpy as np
import pytest
import torch

import oracle
import torch_wcsph as tw
from fluids_amd import scenes
from fluids_amd.adjoint import WCSPHAdjoint

pytestmark = pytest.mark.gpu


def _threads():
    oracle.set_threads(min(16, os.cpu_count() or 1))


def _sorted_rows(a):
    a = np.asarray(a)[:, :3]
    return a[np.lexsort(a.T[::-1])]


def test_adjoint_forward_is_the_solver_trajectory():
    case = scenes.c5_adjoint(nx=12, ny=10, nz=8, steps=30)
    sc = case.scene
    v4 = np.concatenate([case.v0_hidden, np.zeros((sc.n_fluid, 1), np.float32)], 1)
    adj = WCSPHAdjoint(sc.settings, sc.container, sc.volumes[0], v4, v4, checkpoint_every=7)
    adj.forward(30)
    pos, vel, vh = adj.final_state()
    sim = scenes.build(sc)
    sim.set_velocities(v4)
    sim.set_half_velocities(v4)
    sim.step(30)
    assert np.array_equal(_sorted_rows(pos), _sorted_rows(sim.positions()))
    assert np.array_equal(np.sort(vel[:, 0]), np.sort(sim.velocities()[:, 0]))


def _wall_case():
    """A small block next to the +x wall, thrown at it."""
    case = scenes.c5_adjoint(nx=6, ny=8, nz=8, steps=12)
    sc = case.scene
    vol = sc.volumes[0].copy()
    vol[:, 0] += np.float32(0.5 - 0.004 - vol[:, 0].max())
    v = case.v0_hidden.copy()
    v[:, 0] += 3.0
    sc.volumes = [vol]
    return sc, v.astype(np.float32)


@pytest.mark.parametrize("which", ["bulk", "wall"])
def test_adjoint_matches_torch_autograd(which):
    if which == "bulk":
        case = scenes.c5_adjoint(nx=10, ny=10, nz=8, steps=12)
        sc, v0 = case.scene, case.v0_hidden
    else:
        sc, v0 = _wall_case()
    st = sc.settings
    steps = 12
    n = sc.n_fluid
    x0 = sc.volumes[0]
    rng = np.random.default_rng(7)
    wx, wv, wh = [rng.standard_normal((n, 3)) for _ in range(3)]
    # GPU adjoint
    v4 = np.concatenate([v0, np.zeros((n, 1), np.float32)], 1)
    adj = WCSPHAdjoint(st, sc.container, x0, v4, v4, checkpoint_every=5)
    adj.forward(steps)
    g = adj.backward(wx.astype(np.float32), wv.astype(np.float32), wh.astype(np.float32))
    # torch float64 reference
    P = tw.Params(st, sc.container)
    X = torch.tensor(x0[:, :3], dtype=torch.float64, requires_grad=True)
    V = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    W = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    kv = torch.tensor(float(np.float32(st["k_viscosity"])), dtype=torch.float64, requires_grad=True)
    kg = torch.tensor(float(np.float32(st["gas_stiffness"])), dtype=torch.float64, requires_grad=True)
    sg = torch.tensor(float(np.float32(st["surface_tension"])), dtype=torch.float64, requires_grad=True)
    xt, vt, wt = tw.rollout(X, V, W, P, kv, kg, sg, steps)
    loss = (xt * torch.tensor(wx)).sum() + (vt * torch.tensor(wv)).sum() + (wt * torch.tensor(wh)).sum()
    loss.backward()
    if which == "wall":
        assert (xt.detach().numpy()[:, 0] >= 0.5 - 1e-9).any(), "no particle reached the wall"
    pairs = [("pos0", X.grad), ("vel0", V.grad), ("vh0", W.grad)]
    for name, ref in pairs:
        ref = ref.numpy()
        err = np.abs(g[name] - ref).max() / np.abs(ref).max()
        print(f"{which} {name}: max-rel {err:.2e}", flush=True)
        assert err < 2e-3, (which, name, err)
    for name, ref in [("k_viscosity", kv.grad), ("gas_stiffness", kg.grad), ("surface_tension", sg.grad)]:
        ref = float(ref)
        err = abs(g[name] - ref) / max(abs(ref), 1e-30)
        print(f"{which} {name}: {g[name]:.6e} vs {ref:.6e} (rel {err:.2e})", flush=True)
        assert err < 2e-3, (which, name, err)


def _small_case():
    case = scenes.c5_adjoint(nx=8, ny=8, nz=6, steps=12)
    sc = case.scene
    v4 = np.concatenate([case.v0_hidden, np.zeros((sc.n_fluid, 1), np.float32)], 1)
    return case, sc, v4


def test_adjoint_backward_twice_gives_the_same_gradients():
    """Three checkpoint segments: the first backward leaves the tape and the
    stored arrays at segment 0; a second backward must map the cotangent onto
    the same particles and give identical gradients."""
    case, sc, v4 = _small_case()
    n = sc.n_fluid
    adj = WCSPHAdjoint(sc.settings, sc.container, sc.volumes[0], v4, v4, checkpoint_every=5)
    adj.forward(12)
    rng = np.random.default_rng(3)
    gx, gv = rng.standard_normal((2, n, 3)).astype(np.float32)
    a = adj.backward(gx, gv)
    b = adj.backward(gx, gv)
    for k in ("pos0", "vel0", "vh0"):
        assert np.array_equal(a[k], b[k]), k
        assert np.abs(a[k]).max() > 0
    for k in ("k_viscosity", "gas_stiffness", "surface_tension"):
        assert a[k] == b[k], k


def test_rollout_function_two_applies_on_one_handle():
    """Two rollouts through one handle before a single backward: each gets
    its own gradient, equal to that of a rollout run alone; plain-float
    parameters get no gradient and a tensor surface tension gets one."""
    from fluids_amd.adjoint import WCSPHRollout
    case, sc, v4 = _small_case()
    st = sc.settings
    adj = WCSPHAdjoint(st, sc.container, sc.volumes[0], v4, v4, checkpoint_every=5)
    va = torch.tensor(case.v0_hidden, requires_grad=True)
    vb = torch.tensor(0.5 * case.v0_hidden + 0.1, requires_grad=True)
    kv = torch.tensor(float(st["k_viscosity"]), dtype=torch.float64, requires_grad=True)
    xa = WCSPHRollout.apply(va, kv, float(st["gas_stiffness"]), adj, 12)
    xb = WCSPHRollout.apply(vb, kv, float(st["gas_stiffness"]), adj, 12)
    (xa.sum() + (xb ** 2).sum()).backward()
    alone = torch.tensor(case.v0_hidden, requires_grad=True)
    sg = torch.tensor(float(st["surface_tension"]), dtype=torch.float64, requires_grad=True)
    WCSPHRollout.apply(alone, float(st["k_viscosity"]), float(st["gas_stiffness"]), adj, 12, sg).sum().backward()
    assert torch.equal(va.grad, alone.grad)
    assert kv.grad is not None and sg.grad is not None and float(sg.grad) != 0.0
    alone_b = torch.tensor(0.5 * case.v0_hidden + 0.1, requires_grad=True)
    (WCSPHRollout.apply(alone_b, float(st["k_viscosity"]), float(st["gas_stiffness"]), adj, 12) ** 2).sum().backward()
    assert torch.equal(vb.grad, alone_b.grad)


@pytest.fixture(scope="module")
def c5_gradient():
    """Gradient of the C5 loss at a point away from the optimum."""
    _threads()
    case = scenes.c5_adjoint()
    sc = case.scene
    st, n, h = sc.settings, sc.n_fluid, sc.settings["fluid_support_radius"]
    rho0 = st["rest_density"]
    # targets: the hidden run
    v4 = np.concatenate([case.v0_hidden, np.zeros((n, 1), np.float32)], 1)
    adj = WCSPHAdjoint(st, sc.container, sc.volumes[0], v4, v4, checkpoint_every=case.checkpoint_every)
    adj.forward(case.steps)
    target = adj.final_state()[0][:, :3].astype(np.float64)
    # evaluation point
    from fluids_amd.adjoint import WCSPHRollout
    v_eval = (0.7 * case.v0_hidden + 0.05).astype(np.float32)
    alpha_e, c0_e = 0.02, 24.0
    v0 = torch.tensor(v_eval, dtype=torch.float32, requires_grad=True)
    alpha = torch.tensor(alpha_e, dtype=torch.float64, requires_grad=True)
    c0 = torch.tensor(c0_e, dtype=torch.float64, requires_grad=True)
    kv = scenes.monaghan_k_viscosity(alpha, c0, h, case.mass)
    kg = scenes.tait_gas_stiffness(c0, rho0)
    xT = WCSPHRollout.apply(v0, kv, kg, adj, case.steps)
    loss = ((xT.double() - torch.tensor(target)) ** 2).sum(1).mean()
    loss.backward()
    fwd_ms, bwd_ms = adj.timings()
    print(f"C5 loss {loss.item():.6e}; forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms", flush=True)

    def loss64(v, a, c):
        p, _, _, mc = oracle.wcsph_rollout64(st, sc.container, sc.volumes[0], v, v, steps=case.steps,
                                             k_visc=scenes.monaghan_k_viscosity(a, c, h, case.mass),
                                             k_gas=scenes.tait_gas_stiffness(c, rho0))
        assert mc < 50
        return float(np.mean(np.sum((p - target) ** 2, axis=1)))

    return dict(case=case, v_eval=v_eval.astype(np.float64), alpha=alpha_e, c0=c0_e, g_v0=v0.grad.numpy(),
                g_alpha=float(alpha.grad), g_c0=float(c0.grad), loss64=loss64)


def _smooth_direction(x, rng, L):
    d = np.zeros((x.shape[0], 3))
    for c in range(3):
        for _ in range(4):
            k = rng.integers(-2, 3, size=3).astype(np.float64)
            d[:, c] += rng.uniform(-1, 1) * np.sin(2 * np.pi * (x @ k) / L + rng.uniform(0, 2 * np.pi))
    return d / np.sqrt(np.mean(np.sum(d ** 2, axis=1)))  # RMS 1 m/s per particle


def _fd_directions(first, count, g):
    rng = np.random.default_rng(505)
    x = g["case"].scene.volumes[0][:, :3].astype(np.float64)
    dirs = [_smooth_direction(x, rng, 0.32) for _ in range(64)]
    ads, fds = [], []
    eps = 1e-4
    for k in range(first, first + count):
        d = dirs[k]
        fd = (g["loss64"](g["v_eval"] + eps * d, g["alpha"], g["c0"]) -
              g["loss64"](g["v_eval"] - eps * d, g["alpha"], g["c0"])) / (2 * eps)
        ad = float(np.sum(g["g_v0"].astype(np.float64) * d))
        ads.append(ad)
        fds.append(fd)
        print(f"direction {k}: adjoint {ad:.6e} fd {fd:.6e} rel {abs(ad - fd) / abs(fd):.2e}", flush=True)
    ads, fds = np.array(ads), np.array(fds)
    scale = np.sqrt(np.mean(fds ** 2))
    err_scaled = np.abs(ads - fds) / scale
    rel = np.abs(ads - fds) / np.abs(fds)
    print(f"directions {first}-{first + count - 1}: max |ad-fd|/rms {err_scaled.max():.2e}, "
          f"median rel {np.median(rel):.2e}", flush=True)
    return err_scaled, rel


def test_c5_gradient_fd_directions_0_31_and_alpha(c5_gradient):
    g = c5_gradient
    err_scaled, rel = _fd_directions(0, 32, g)
    ea = 1e-4 * g["alpha"]
    fd = (g["loss64"](g["v_eval"], g["alpha"] + ea, g["c0"]) - g["loss64"](g["v_eval"], g["alpha"] - ea, g["c0"])) / (2 * ea)
    err_a = abs(g["g_alpha"] - fd) / abs(fd)
    print(f"alpha: adjoint {g['g_alpha']:.6e} fd {fd:.6e} rel {err_a:.2e}", flush=True)
    assert err_scaled.max() < 1e-2
    assert np.median(rel) < 5e-3
    assert err_a < 1e-2


def test_c5_gradient_fd_directions_32_63_and_c0(c5_gradient):
    g = c5_gradient
    err_scaled, rel = _fd_directions(32, 32, g)
    ec = 1e-4 * g["c0"]
    fd = (g["loss64"](g["v_eval"], g["alpha"], g["c0"] + ec) - g["loss64"](g["v_eval"], g["alpha"], g["c0"] - ec)) / (2 * ec)
    err_c = abs(g["g_c0"] - fd) / abs(fd)
    print(f"c0: adjoint {g['g_c0']:.6e} fd {fd:.6e} rel {err_c:.2e}", flush=True)
    assert err_scaled.max() < 1e-2
    assert np.median(rel) < 5e-3
    assert err_c < 1e-2
