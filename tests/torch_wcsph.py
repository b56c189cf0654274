"""A differentiable float64 restatement of the WCSPH step in torch, used only
by the tests as the reference gradient of the hand-written adjoint: torch's
autograd derives the gradient of the same formulas independently.

It follows wcsph.cl:32-224 (density with the repaired W_POLY6 on r^2 when
the flag is set, Tait pressure, normals, pressure / viscosity / surface
tension accelerations, the leap-frog with the reference's double dt*a, the
velocity clamps and the clamp-box reflection). Neighbour pairs are found by
brute force each step (r^2 below the squared 6-decimal support literal), so
it is meant for small scenes.
"""
import math

import numpy as np
import torch

import oracle


class Params:
    def __init__(self, settings: dict, container):
        h = np.float32(settings["fluid_support_radius"])
        r = np.float32(settings["fluid_particle_radius"])
        rho0 = np.float32(settings["rest_density"])
        self.h = float(h)
        self.mass = float(rho0 * np.float32(2.0 * r) ** 3)
        self.rho0 = float(rho0)
        # the host computes each constant in double and passes it as a float
        # (wcsphsimluation.cpp:188-199)
        f32 = lambda v: float(np.float32(v))
        self.poly6_eval = f32(315.0 / (64.0 * math.pi * self.h ** 9))
        self.poly6_grad = f32(-945.0 / (32.0 * math.pi * self.h ** 9))
        self.spiky_grad = f32(-45.0 / (math.pi * self.h ** 6))
        self.visc_lapl = f32(45.0 / (math.pi * self.h ** 6))
        self.st_main = f32(32.0 / (math.pi * float(np.float32(h) ** np.float32(9))))
        self.st_term = f32(np.float32(h) ** np.float32(6) / np.float32(64.0))
        lit = float(np.float32(oracle.define_literal(h)))
        self.h_d = lit
        self.cut2 = float(np.float32(lit) * np.float32(lit))
        self.mass_d = float(np.float32(oracle.define_literal(np.float32(self.mass))))
        self.dt = float(np.float32(settings["time_step"]))
        self.vmax = float(np.float32(settings["max_vel"]))
        self.g = -float(np.float32(settings["gravity"]))
        self.fix = bool(settings["flags"] & 1)
        W, H, D = [float(np.float32(v)) for v in container]
        self.lo = [-W / 2, -H / 2, -D / 2]
        self.hi = [W / 2, 10.0, D / 2]


def _pairs(x, cut2):
    with torch.no_grad():
        d2 = torch.cdist(x, x) ** 2
        m = d2 < cut2
        m.fill_diagonal_(False)
        I, J = torch.nonzero(m, as_tuple=True)
    return I, J


def step(x, v, vh, P: Params, k_visc, k_gas, sigma):
    n = x.shape[0]
    I, J = _pairs(x, P.cut2)
    r = x[I] - x[J]
    s2 = (r * r).sum(1)
    q = torch.sqrt(s2)
    h2 = P.h * P.h
    t = (h2 - (s2 if P.fix else q)) ** 3
    dens = torch.full((n,), h2 ** 3, dtype=x.dtype).index_add(0, I, t)
    rho = dens * (P.poly6_eval * P.mass)
    p = torch.clamp(k_gas * ((rho / P.rho0) ** 7 - 1.0), min=0.0)
    w = P.h_d * P.h_d - s2
    nrm = torch.zeros_like(x).index_add(0, I, r * (w * w / rho[J])[:, None]) * (P.poly6_grad * P.mass_d * P.h_d)
    ok = (rho[I] > 0) & (rho[J] > 0) & (q > 1e-5)
    I, J, r, q = I[ok], J[ok], r[ok], q[ok]
    hq = P.h - q
    C = p / rho ** 2
    pacc = torch.zeros_like(x).index_add(0, I, r * (P.spiky_grad / q * hq ** 2 * (C[I] + p[J] / rho[J] ** 2))[:, None])
    vacc = torch.zeros_like(x).index_add(0, I, (v[J] - v[I]) * (hq / rho[J])[:, None])
    K = 2.0 * P.rho0 / (rho[I] + rho[J])
    fcurv = torch.zeros_like(x).index_add(0, I, (nrm[I] - nrm[J]) * K[:, None])
    st = torch.where(q <= 0.5 * P.h, 2.0 * hq ** 3 * q ** 3 - P.st_term,
                     torch.where(q <= P.h, hq ** 3 * q ** 3, torch.zeros_like(q)))
    fcoh = torch.zeros_like(x).index_add(0, I, r * (st / q * K)[:, None])
    g = torch.tensor([0.0, P.g, 0.0], dtype=x.dtype)
    a = g - sigma * (fcurv + P.st_main * fcoh) - P.mass * pacc + (P.visc_lapl * k_visc) * vacc
    dt, vmax = P.dt, P.vmax
    w1 = vh + dt * a
    xp = x + dt * torch.clamp(w1 + dt * a, -vmax, vmax)
    lo = torch.tensor(P.lo, dtype=x.dtype)
    hi = torch.tensor(P.hi, dtype=x.dtype)
    cp = torch.maximum(torch.minimum(xp, hi), lo)
    dv = cp - xp
    dd2 = (dv * dv).sum(1)
    vn2 = (w1 * w1).sum(1)
    hit = (dd2 > 0) & (vn2 > 0)
    dd = torch.sqrt(torch.where(hit, dd2, torch.ones_like(dd2)))
    vn = torch.sqrt(torch.where(hit, vn2, torch.ones_like(vn2)))
    with torch.no_grad():
        sg = torch.sign(dv)
        nn = sg / torch.sqrt(torch.clamp((sg * sg).sum(1, keepdim=True), min=1.0))
    rest = dd / (dt * vn)
    w2 = w1 - nn * ((1.0 + rest) * (w1 * nn).sum(1))[:, None]
    w1 = torch.where(hit[:, None], w2, w1)
    xp = torch.where(hit[:, None], cp, xp)
    v_out = torch.clamp(w1 + 0.5 * dt * a, -vmax, vmax)
    vh_out = torch.clamp(w1, -vmax, vmax)
    return xp, v_out, vh_out


def rollout(x0, v0, vh0, P: Params, k_visc, k_gas, sigma, steps):
    x, v, vh = x0, v0, vh0
    for _ in range(steps):
        x, v, vh = step(x, v, vh, P, k_visc, k_gas, sigma)
    return x, v, vh
