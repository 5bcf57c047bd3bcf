"""The FP64 torch restatement used as the gradient checker (oracle/rollout64)
agrees with the C oracle (forces, velocities, Euler step) and with
closed-form / finite-difference checks of its loss."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import rollout64 as r64


def _system(n=120, seed=3, L=4e-4):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, L, size=(n, 3))
    radii = np.exp(rng.uniform(np.log(2e-6), np.log(1.2e-5), size=n))
    pos[1] = pos[0] + np.array([0.0, 0.0, 1.0e-5])  # one overlapping pair (masked)
    pos %= L
    return pos, radii, L


def test_velocity_matches_c_oracle():
    pos, radii, L = _system()
    E, mu, mu_w, eps = 3e5, 0.05, 0.001, 2.85
    m, eta, sf = r64.m_const(eps, E), r64.eta_const(mu), r64.stokes_factor(mu, mu_w)
    v, F, V = r64.velocity(torch.tensor(pos), torch.tensor(radii), m, eta, sf, L,
                           return_parts=True)
    F_ref = oracle.direct_forces(pos, radii, m, L, True)
    V_ref = oracle.direct_convection(pos, F_ref, eta, L, True)
    np.testing.assert_allclose(F.numpy(), F_ref, rtol=1e-11, atol=1e-11 * np.abs(F_ref).max())
    np.testing.assert_allclose(V.numpy(), V_ref, rtol=1e-10, atol=1e-10 * np.abs(V_ref).max())
    np.testing.assert_allclose(m, oracle.m_const(eps, E), rtol=1e-15)
    np.testing.assert_allclose(sf, oracle.stokes_factor(mu, mu_w), rtol=1e-15)
    np.testing.assert_allclose(v.numpy(), F_ref / (sf * radii[:, None]) + V_ref, rtol=1e-10,
                               atol=1e-10 * np.abs(V_ref).max())


def test_rollout_matches_c_oracle_steps():
    pos, radii, L = _system(n=60, seed=5)
    E, mu, mu_w, eps = 4e5, 0.05, 0.001, 2.85
    m, eta, sf = r64.m_const(eps, E), r64.eta_const(mu), r64.stokes_factor(mu, mu_w)
    dt = 1e-3
    x = pos.copy()
    for _ in range(3):
        F = oracle.direct_forces(x, radii, m, L, True)
        V = oracle.direct_convection(x, F, eta, L, True)
        x = oracle.euler_step(x, F, V, radii, sf, dt, L, True)
    got = r64.rollout(torch.tensor(pos), torch.tensor(radii), E, mu, eps, mu_w, dt, 3, L)
    d = got.numpy() - x
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-12 * L


def test_contact_loss_against_a_loop_and_fd():
    pos, radii, L = _system(n=25, seed=8, L=6e-5)
    x = torch.tensor(pos, requires_grad=True)
    loss = r64.contact_loss(x, torch.tensor(radii), L)
    ref = 0.0
    for i in range(len(radii)):
        for j in range(i + 1, len(radii)):
            d = pos[j] - pos[i]
            d -= L * np.round(d / L)
            s = 1e6 * (1.05 * (radii[i] + radii[j]) - np.linalg.norm(d))
            ref += np.log1p(np.exp(-abs(s))) + max(s, 0.0)
    assert abs(loss.item() - ref) <= 1e-12 * abs(ref)
    loss.backward()
    g = x.grad.numpy()
    h = 1e-12
    for k in (0, 7, 19):
        for a in range(3):
            xp = pos.copy()
            xp[k, a] += h
            xm = pos.copy()
            xm[k, a] -= h
            fd = (r64.contact_loss(torch.tensor(xp), torch.tensor(radii), L).item()
                  - r64.contact_loss(torch.tensor(xm), torch.tensor(radii), L).item()) / (2 * h)
            assert fd == pytest.approx(g[k, a], rel=1e-5, abs=1e-3 * np.abs(g).max())
