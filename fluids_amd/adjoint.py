"""Gradients through WCSPH rollouts (include/sph.h, sph_adjoint_*): a
hand-written adjoint of the solver's step, with checkpoints, exposed as a
torch.autograd.Function.

    adj = WCSPHAdjoint(settings, container, pos0, checkpoint_every=20)
    x_T = WCSPHRollout.apply(v0, k_visc, k_gas, adj, steps)   # (n, 3)
    loss = ((x_T - target) ** 2).sum(1).mean()
    loss.backward()          # v0.grad, k_visc.grad, k_gas.grad

A sixth argument gives the surface tension coefficient (a float, or a tensor
that then receives its gradient); k_visc and k_gas may be plain floats, whose
gradients are then not formed. Several rollouts may share one handle: each
backward rebuilds its own rollout's tape if another forward ran since.

v0 sets both the velocity and the half-step velocity of the initial state,
so its gradient is the sum of the two. The forward pass is the solver's
(bit for bit the same trajectory as sph_step); the backward pass runs on the
GPU through the C ABI and has no CPU fallback.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib
from .batch import _settings


def _f4(a, n):
    if a is None:
        return None
    a = np.asarray(a, np.float32)
    if a.shape == (n, 3):
        a = np.concatenate([a, np.zeros((n, 1), np.float32)], axis=1)
    a = np.ascontiguousarray(a.reshape(n, 4))
    return a


def _p(a):
    return None if a is None else a.ctypes.data


class WCSPHAdjoint:
    def __init__(self, settings, container, pos0, vel0=None, vh0=None, checkpoint_every=20):
        pos0 = np.asarray(pos0, np.float32)
        self.n = pos0.shape[0]
        p = _f4(pos0[:, :3], self.n)
        p[:, 3] = 1.0
        self.settings = dict(settings) if isinstance(settings, dict) else None
        st = _settings(settings)
        v, w = _f4(vel0, self.n), _f4(vh0, self.n)
        self._h = lib().sph_adjoint_create(C.byref(st), float(container[0]), float(container[1]),
                                           float(container[2]), p.ctypes.data, _p(v), _p(w), self.n,
                                           int(checkpoint_every))
        if not self._h:
            raise RuntimeError(lib().sph_last_error().decode())
        self._pos0 = p

    def close(self):
        if getattr(self, "_h", None):
            lib().sph_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_initial(self, pos=None, vel=None, vh=None):
        p = self._pos0 if pos is None else _f4(np.asarray(pos)[:, :3], self.n)
        p[:, 3] = 1.0
        v, w = _f4(vel, self.n), _f4(vh, self.n)
        check(lib().sph_adjoint_set_initial(self._h, p.ctypes.data, _p(v), _p(w)))

    def set_params(self, k_viscosity, gas_stiffness, surface_tension):
        check(lib().sph_adjoint_set_params(self._h, float(k_viscosity), float(gas_stiffness),
                                           float(surface_tension)))

    def forward(self, steps):
        check(lib().sph_adjoint_forward(self._h, int(steps)))
        # the handle holds one tape: a rollout whose backward comes after
        # another forward on the same handle must run its own forward again
        self.generation = getattr(self, "generation", 0) + 1

    def final_state(self):
        out = [np.zeros((self.n, 4), np.float32) for _ in range(3)]
        check(lib().sph_adjoint_final_state(self._h, *[o.ctypes.data for o in out]))
        return out

    def backward(self, g_pos=None, g_vel=None, g_vh=None) -> dict:
        gp, gv, gw = _f4(g_pos, self.n), _f4(g_vel, self.n), _f4(g_vh, self.n)
        out = [np.zeros((self.n, 4), np.float32) for _ in range(3)]
        prm = (C.c_double * 3)()
        check(lib().sph_adjoint_backward(self._h, _p(gp), _p(gv), _p(gw), *[o.ctypes.data for o in out], prm))
        return {"pos0": out[0][:, :3], "vel0": out[1][:, :3], "vh0": out[2][:, :3],
                "k_viscosity": prm[0], "gas_stiffness": prm[1], "surface_tension": prm[2]}

    def timings(self):
        f, b = C.c_double(), C.c_double()
        check(lib().sph_adjoint_timings(self._h, C.byref(f), C.byref(b)))
        return f.value, b.value


def _torch_function():
    import torch

    class WCSPHRollout(torch.autograd.Function):
        """x_T = rollout(v0, k_viscosity, gas_stiffness): final positions
        (n, 3) after `steps` WCSPH steps from the adjoint's initial positions
        with initial velocity and half-step velocity v0."""

        @staticmethod
        def forward(ctx, v0, k_visc, k_gas, adj, steps, surface_tension=None):
            v = v0.detach().cpu().numpy().astype(np.float32)
            sigma = adj.settings["surface_tension"] if surface_tension is None else float(surface_tension)
            params = (float(k_visc), float(k_gas), sigma)
            adj.set_initial(None, v, v)
            adj.set_params(*params)
            adj.forward(int(steps))
            pos, _, _ = adj.final_state()
            ctx.adj = adj
            ctx.replay = (v, params, int(steps), adj.generation)
            ctx.meta = (v0.dtype, v0.device)
            ctx.pmeta = tuple((a.dtype, a.device) if torch.is_tensor(a) else None
                              for a in (k_visc, k_gas, surface_tension))
            return torch.from_numpy(pos[:, :3].copy()).to(device=v0.device, dtype=v0.dtype)

        @staticmethod
        def backward(ctx, g_pos):
            adj = ctx.adj
            v, params, steps, gen = ctx.replay
            if adj.generation != gen:
                # another rollout ran forward on this handle since: rebuild
                # this rollout's tape (same inputs, same bits) first
                adj.set_initial(None, v, v)
                adj.set_params(*params)
                adj.forward(steps)
            g = adj.backward(g_pos.detach().cpu().numpy().astype(np.float32))
            dtype, device = ctx.meta
            g_v0 = torch.from_numpy((g["vel0"] + g["vh0"]).astype(np.float32)).to(device=device, dtype=dtype)
            out = []
            for key, meta in zip(("k_viscosity", "gas_stiffness", "surface_tension"), ctx.pmeta):
                out.append(None if meta is None else torch.tensor(g[key], dtype=meta[0], device=meta[1]))
            return g_v0, out[0], out[1], None, None, out[2]

    return WCSPHRollout


_ROLLOUT = None


def __getattr__(name):
    # the torch class is built on first use, so importing this module does
    # not import torch
    global _ROLLOUT
    if name == "WCSPHRollout":
        if _ROLLOUT is None:
            _ROLLOUT = _torch_function()
        return _ROLLOUT
    raise AttributeError(name)
