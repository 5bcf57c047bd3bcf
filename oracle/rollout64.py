"""Float64 torch restatement of the direct DEM step, rollout and contact loss —
TEST INFRASTRUCTURE ONLY (the gradient checker).

Only ``tests/`` and ``bench.py`` import this module, as the reference the
hand-written adjoint is checked against: autograd through this plain FP64
path, and central finite differences of it.  The product never imports it.

* ``velocity(x, radii, m, eta, sf, L)`` — one evaluation of the direct path's
  Euler velocity v = F / (sf r) + V:
  F per ``_calculate_forces`` / ``_calculate_one_force``
  (dem/force_calculator/direct_droplet_force_calculator.py:183-253:
  d = x_j - x_i, minimum image |d| > L/2 -> d -= sign(d) L, mask
  |d| > r_i + r_j), V per ``_calculate_convection_velocities`` (:354-411: all
  j != i, no mask, V_i = eta sum_j F_j/|x| + x (x.F_j)/|x|^3), sf the Stokes
  factor (dem/solver/euler_droplet_solver.py:37-38, :124-125).
* ``rollout`` — T explicit Euler steps x += v dt, x %= L (:113-131).
* ``contact_loss`` — sum over pairs i < j of
  softplus(1.05 (r_i + r_j) - |x_i - x_j|) with lengths in micrometres
  (minimum image): a smooth count of near-contacts.
* ``m_const(eps_oil, E)`` = 12 pi eps0 eps_oil E^2 (:30), ``eta_const(mu)`` =
  1/(8 pi mu) (:31), ``stokes_factor(mu, mu_w)`` = 2 pi mu (2 mu + 3 mu_w) /
  (mu + mu_w).

Pinned (tests/test_oracle_rollout.py) against the C oracle's forces,
velocities and Euler step on the same inputs.  Dense N x N: O(N^2) memory,
meant for the GPU at N = 4096 or the CPU at small N.
"""

from __future__ import annotations

import math

import torch

EPS0 = 8.85418781762039e-12


def m_const(eps_oil, E):
    return 12 * math.pi * EPS0 * eps_oil * E**2


def eta_const(mu):
    return 1 / (8 * math.pi * mu)


def stokes_factor(mu, mu_w):
    return 2 * math.pi * mu * (2 * mu + 3 * mu_w) / (mu + mu_w)


def _mic(d, L):
    return torch.where(d.abs() > L / 2, d - torch.sign(d) * L, d)


def velocity(x, radii, m, eta, sf, L, return_parts: bool = False):
    """v = F / (sf r) + V for positions x (N,3); m, eta, sf may be tensors."""
    n = x.shape[0]
    d = _mic(x[None, :, :] - x[:, None, :], L)  # d[i, j] = x_j - x_i
    eye = torch.eye(n, dtype=torch.bool, device=x.device)
    r2 = (d * d).sum(-1)
    r = torch.sqrt(torch.where(eye, torch.ones_like(r2), r2))
    dx, dy, dz = d[..., 0], d[..., 1], d[..., 2]
    q = dx * dx + dy * dy
    a_xy = 4 * dz * dz - q
    a_z = 2 * dz * dz - 3 * q
    r3 = radii**3
    keep = (r > radii[:, None] + radii[None, :]) & ~eye
    w = torch.where(keep, r3[:, None] * r3[None, :] / r**7, torch.zeros_like(r))
    F = m * torch.stack([(w * dx * a_xy).sum(1), (w * dy * a_xy).sum(1),
                         (w * dz * a_z).sum(1)], -1)
    inv = torch.where(eye, torch.zeros_like(r), 1 / r)
    dF = (d * F[None, :, :]).sum(-1)  # d_ij . F_j
    V = eta * ((inv[..., None] * F[None, :, :]).sum(1)
               + ((inv**3 * dF)[..., None] * d).sum(1))
    v = F / (sf * radii[:, None]) + V
    if return_parts:
        return v, F, V
    return v


def rollout(x0, radii, E, mu, eps_oil, mu_w, dt, steps, L, checkpoint_every: int = 0):
    """x_T of T Euler steps from x0 (differentiable in x0, E, mu)."""
    m = m_const(eps_oil, E)
    eta = eta_const(mu)
    sf = stokes_factor(mu, mu_w)

    def seg(x, k):
        for _ in range(k):
            x = torch.remainder(x + velocity(x, radii, m, eta, sf, L) * dt, L)
        return x

    x = x0
    if checkpoint_every and x0.requires_grad:
        from torch.utils.checkpoint import checkpoint

        t = 0
        while t < steps:
            k = min(checkpoint_every, steps - t)
            x = checkpoint(seg, x, k, use_reentrant=False)
            t += k
        return x
    return seg(x, steps)


def contact_loss(x, radii, L, factor: float = 1.05, unit: float = 1e6):
    """sum_{i<j} softplus(factor (r_i + r_j) - |d_ij|), lengths times `unit`."""
    n = x.shape[0]
    d = _mic(x[None, :, :] - x[:, None, :], L)
    eye = torch.eye(n, dtype=torch.bool, device=x.device)
    r2 = (d * d).sum(-1)
    r = torch.sqrt(torch.where(eye, torch.ones_like(r2), r2))
    s = unit * (factor * (radii[:, None] + radii[None, :]) - r)
    sp = torch.nn.functional.softplus(s)
    upper = torch.triu(torch.ones(n, n, dtype=torch.bool, device=x.device), 1)
    return torch.where(upper, sp, torch.zeros_like(sp)).sum()


def central_difference(x0, radii, E, mu, eps_oil, mu_w, dt, steps, L, direction=None,
                       eps=0.0, dE=0.0, dmu=0.0):
    """(loss(+) - loss(-)) / 2 in float64 for the perturbed inputs
    (x0 +- eps v, E +- dE, mu +- dmu), and whether the two trajectories keep
    identical contact masks and periodic images of every pair at every step
    (the loss jumps where either switches, so a difference quotient across a
    switch is not a derivative)."""
    xp = x0 + eps * direction if direction is not None else x0.clone()
    xm = x0 - eps * direction if direction is not None else x0.clone()
    par = []
    for sgn in (1.0, -1.0):
        Ev, muv = E + sgn * dE, mu + sgn * dmu
        par.append((m_const(eps_oil, Ev), eta_const(muv), stokes_factor(muv, mu_w)))
    thr = radii[:, None] + radii[None, :]
    same = True
    with torch.no_grad():
        for _ in range(steps):
            if same:
                da = xp[None, :, :] - xp[:, None, :]
                db = xm[None, :, :] - xm[:, None, :]
                if not torch.equal(da.abs() > L / 2, db.abs() > L / 2):
                    same = False
                elif not torch.equal((_mic(da, L) ** 2).sum(-1).sqrt() > thr,
                                     (_mic(db, L) ** 2).sum(-1).sqrt() > thr):
                    same = False
                del da, db
            xp = torch.remainder(xp + velocity(xp, radii, *par[0], L) * dt, L)
            xm = torch.remainder(xm + velocity(xm, radii, *par[1], L) * dt, L)
        q = (contact_loss(xp, radii, L) - contact_loss(xm, radii, L)).item() / 2
    return q, same


def switch_free_derivative(x0, radii, E, mu, eps_oil, mu_w, dt, steps, L, which,
                           direction=None,
                           hs=(1e-6, 1e-7, 1e-8, 1e-9, 1e-10, 1e-11, 1e-12)):
    """Central difference of the loss along `which` ("E", "mu" or "x0" with
    `direction`), with the largest relative step h in `hs` (positions: h L)
    whose two trajectories switch no contact mask or periodic image.
    Returns (derivative, h) or (None, None) if every step straddles a switch."""
    for h in hs:
        if which == "E":
            q, same = central_difference(x0, radii, E, mu, eps_oil, mu_w, dt, steps, L,
                                         dE=h * E)
            step = h * E
        elif which == "mu":
            q, same = central_difference(x0, radii, E, mu, eps_oil, mu_w, dt, steps, L,
                                         dmu=h * mu)
            step = h * mu
        else:
            q, same = central_difference(x0, radii, E, mu, eps_oil, mu_w, dt, steps, L,
                                         direction=direction, eps=h * L)
            step = h * L
        if same:
            return q / step, h
    return None, None
