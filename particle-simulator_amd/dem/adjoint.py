"""Differentiable direct DEM trajectories with a hand-written adjoint.

The reference integrates (dem/solver/euler_droplet_solver.py:113-131)
    x_{t+1} = (x_t + dt v(x_t)) mod L,   v = F / (sf r) + V(x, F),
with the dipole forces F (direct_droplet_force_calculator.py:183-253) and the
Oseen velocities V (:354-411); it has no gradients.  Here a trajectory of T
steps (no merging) is a ``torch.autograd.Function`` of the initial positions
x0, the field strength E (m_const = 12 pi eps0 eps_oil E^2) and the oil
viscosity mu (eta_const = 1/(8 pi mu) and the Stokes factor
sf = 2 pi mu (2 mu + 3 mu_w)/(mu + mu_w)):

* forward: the FP32 direct kernels and the float64 Euler kernel, keeping the
  state every ``checkpoint_every`` (20) steps;
* backward: the discrete adjoint lambda_t = lambda_{t+1} + dt (dv/dx_t)^T
  lambda_{t+1}, lambda_T = dL/dx_T, recomputing each 20-step segment from
  its checkpoint; (dv/dx)^T lambda and the parameter partials come from
  ``dem_direct_step_vjp`` (the adjoint HIP kernels), so
    dL/dx0 = lambda_0,
    dL/dE  = sum_t dt (2 / E) S_E,t,
    dL/dmu = sum_t dt (-S_eta,t / mu - S_sf,t sf'(mu) / sf).

``contact_loss`` is the near-contact objective sum_{i<j} softplus(1.05
(r_i + r_j) - |x_i - x_j|) with lengths in micrometres (minimum image), with
its gradient from the same library (``dem_contact_loss``).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _device, _native

EPS0 = 8.85418781762039e-12


def _m_const(eps_oil, E):
    return 12 * math.pi * EPS0 * eps_oil * E**2


def _eta_const(mu):
    return 1 / (8 * math.pi * mu)


def _stokes(mu, mu_w):
    return 2 * math.pi * mu * (2 * mu + 3 * mu_w) / (mu + mu_w)


def _stokes_dmu(mu, mu_w):
    n = mu * (2 * mu + 3 * mu_w)
    d = mu + mu_w
    return 2 * math.pi * ((4 * mu + 3 * mu_w) * d - n) / (d * d)


class DirectTrajectory:
    """T explicit Euler steps of the direct periodic path, differentiable."""

    def __init__(self, radii, L: float, dt: float, steps: int, eps_oil: float = 2.85,
                 eta_water: float = 0.001, checkpoint_every: int = 20):
        self.radii = _device.to_device(np.asarray(radii, dtype=np.float64))
        self.n = int(self.radii.shape[0])
        self.L = float(L)
        self.dt = float(dt)
        self.steps = int(steps)
        self.eps_oil = float(eps_oil)
        self.eta_water = float(eta_water)
        self.checkpoint_every = max(1, int(checkpoint_every))
        self._ws_fwd = _device.Workspace()
        self._ws_adj = _device.Workspace()
        self.recomputed_steps = 0

    def _phys(self, E, mu):
        return _native.DemPhys(_m_const(self.eps_oil, E), _eta_const(mu), self.L, 1, 0)

    def _step(self, x, ph, sf, F, V):
        """x <- x + dt v(x) (in place); F, V receive the forces and velocities."""
        lib = _native.load()
        ws = self._ws_fwd.get(lib.dem_direct_workspace_bytes(self.n))
        st = _device.stream()
        _native.call("dem_direct_forces_and_convection", _device.ptr(x), _device.ptr(self.radii),
                     self.n, ctypes.byref(ph), None, _device.ptr(F), _device.ptr(V),
                     _device.ptr(ws), ws.numel(), st)
        _native.call("dem_euler_step", _device.ptr(x), _device.ptr(F), _device.ptr(V),
                     _device.ptr(self.radii), self.n, sf, self.dt, self.L, 1, st)

    def run(self, x0, E, mu):
        """x_T and the checkpoints {t: x_t} (t multiple of checkpoint_every)."""
        x = x0.detach().to(device=_device.device(), dtype=_device.torch().float64).contiguous().clone()
        ph = self._phys(E, mu)
        sf = _stokes(mu, self.eta_water)
        F = _device.empty((self.n, 3))
        V = _device.empty((self.n, 3))
        ckpt = {}
        for t in range(self.steps):
            if t % self.checkpoint_every == 0:
                ckpt[t] = x.clone()
            self._step(x, ph, sf, F, V)
        return x, ckpt

    def adjoint(self, ckpt, lam_T, E, mu):
        """(dL/dx0, dL/dE, dL/dmu) for dL/dx_T = lam_T."""
        t_ = _device.torch()
        lib = _native.load()
        ph = self._phys(E, mu)
        sf = _stokes(mu, self.eta_water)
        lam = lam_T.detach().to(dtype=t_.float64).contiguous().clone()
        gx = _device.empty((self.n, 3))
        sums = t_.zeros((self.steps, 3), dtype=t_.float64, device=lam.device)
        ws = self._ws_adj.get(lib.dem_adjoint_workspace_bytes(self.n))
        st = _device.stream()
        starts = sorted(ckpt)
        for k in reversed(range(len(starts))):
            t0 = starts[k]
            t1 = starts[k + 1] if k + 1 < len(starts) else self.steps
            # recompute the segment's states and forces from its checkpoint
            xs, Fs = [], []
            x = ckpt[t0].clone()
            V = _device.empty((self.n, 3))
            for _ in range(t0, t1):
                xs.append(x.clone())
                F = _device.empty((self.n, 3))
                self._step(x, ph, sf, F, V)
                Fs.append(F)
                self.recomputed_steps += 1
            for t in reversed(range(t0, t1)):
                _native.call("dem_direct_step_vjp", _device.ptr(xs[t - t0]),
                             _device.ptr(self.radii), _device.ptr(Fs[t - t0]), _device.ptr(lam),
                             self.n, ctypes.byref(ph), sf, _device.ptr(gx),
                             _device.ptr(sums[t]), _device.ptr(ws), ws.numel(), st)
                lam.add_(gx, alpha=self.dt)
        S = sums.sum(0).cpu().numpy()
        dE = self.dt * 2.0 / E * S[0]
        dmu = self.dt * (-S[1] / mu - S[2] * _stokes_dmu(mu, self.eta_water) / sf)
        return lam, float(dE), float(dmu)


def _autograd_function():
    torch = _device.torch()

    class _Rollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x0, E, mu, traj):
            Ev, muv = float(E), float(mu)
            xT, ckpt = traj.run(x0, Ev, muv)
            ctx.traj, ctx.ckpt, ctx.params = traj, ckpt, (Ev, muv)
            ctx.like = (E, mu)
            return xT

        @staticmethod
        def backward(ctx, g):
            Ev, muv = ctx.params
            gx0, gE, gmu = ctx.traj.adjoint(ctx.ckpt, g, Ev, muv)
            E, mu = ctx.like
            tE = torch.as_tensor(gE, dtype=torch.float64) if torch.is_tensor(E) else None
            tmu = torch.as_tensor(gmu, dtype=torch.float64) if torch.is_tensor(mu) else None
            return gx0, tE, tmu, None

    class _ContactLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, traj, factor, unit):
            lib = _native.load()
            n = traj.n
            ph = _native.DemPhys(1.0, 1.0, traj.L, 1, 0)
            loss = torch.zeros((), dtype=torch.float64, device=_device.device())
            grad = _device.empty((n, 3))
            ws = traj._ws_adj.get(lib.dem_adjoint_workspace_bytes(n))
            xc = x.detach().to(dtype=torch.float64).contiguous()
            _native.call("dem_contact_loss", _device.ptr(xc), _device.ptr(traj.radii), n,
                         ctypes.byref(ph), float(factor), float(unit), _device.ptr(loss),
                         _device.ptr(grad), _device.ptr(ws), ws.numel(), _device.stream())
            ctx.save_for_backward(grad)
            return loss

        @staticmethod
        def backward(ctx, g):
            (grad,) = ctx.saved_tensors
            return grad * g, None, None, None

    return _Rollout, _ContactLoss


_FNS = None


def _fns():
    global _FNS
    if _FNS is None:
        _FNS = _autograd_function()
    return _FNS


def rollout(x0, E, mu, traj: DirectTrajectory):
    """x_T of ``traj`` from x0 (device float64 (N,3)); differentiable in x0 and,
    when given as tensors, in E and mu."""
    return _fns()[0].apply(x0, E, mu, traj)


def contact_loss(x, traj: DirectTrajectory, factor: float = 1.05, unit: float = 1e6):
    """sum_{i<j} softplus(unit (factor (r_i + r_j) - |d_ij|)), differentiable in x."""
    return _fns()[1].apply(x, traj, factor, unit)


__all__ = ["DirectTrajectory", "rollout", "contact_loss"]
