"""Hand-written adjoint (dem.adjoint, ``dem_direct_step_vjp`` /
``dem_contact_loss``) against the FP64 torch restatement (oracle/rollout64):

* one step's vector-Jacobian product (dv/dx)^T a and the parameter partials
  against autograd through the FP64 step;
* the near-contact loss and its gradient;
* gradients of a trajectory (x0, E, mu) through T steps with checkpointed
  recomputation, against autograd through the plain FP64 path and against
  FP64 central finite differences (h = 1e-6 relative).

Tolerance 1e-4 relative for trajectory gradients; 1e-5 (field norm) for the
single-step products, whose FP32 pair sums carry the forward kernels'
accuracy.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from dem import _native
from dem.adjoint import DirectTrajectory, contact_loss, rollout
from dem.force_calculator import DirectDropletForceCalculator
from dem.workloads import differentiable, differentiable_dt, make_system
from oracle import rollout64 as r64

pytestmark = pytest.mark.gpu

MU_W = 0.001
EPS = 2.85


def _field_err(x, ref):
    x = np.asarray(x).reshape(-1)
    ref = np.asarray(ref).reshape(-1)
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def _rel(x, ref):
    return abs(x - ref) / abs(ref)


def test_step_vjp_matches_fp64_autograd(gpu):
    s = make_system(700, seed=2, phi=0.04)
    dev = torch.device("cuda")
    E, mu = 3.5e5, 0.05
    x = torch.tensor(s.positions, device=dev, requires_grad=True)
    r = torch.tensor(s.radii, device=dev)
    m = torch.tensor(r64.m_const(EPS, E), dtype=torch.float64, device=dev, requires_grad=True)
    eta = torch.tensor(r64.eta_const(mu), dtype=torch.float64, device=dev, requires_grad=True)
    sf = torch.tensor(r64.stokes_factor(mu, MU_W), dtype=torch.float64, device=dev,
                      requires_grad=True)
    v = r64.velocity(x, r, m, eta, sf, s.L)
    a = torch.randn(s.n, 3, dtype=torch.float64, device=dev, generator=torch.Generator(dev).manual_seed(1)) * 1e6
    (a * v).sum().backward()

    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=EPS, eta_oil=mu,
                                      eta_water=MU_W, E=E, L=s.L)
    xd = x.detach().contiguous()
    F, _ = fc.forces_and_convection_device(xd, r)
    lib = _native.load()
    ws = torch.empty(lib.dem_adjoint_workspace_bytes(s.n), dtype=torch.uint8, device=dev)
    gx = torch.empty(s.n, 3, dtype=torch.float64, device=dev)
    sums = torch.empty(3, dtype=torch.float64, device=dev)
    import ctypes

    ph = _native.DemPhys(m.item(), eta.item(), s.L, 1, 0)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    _native.call("dem_direct_step_vjp", p(xd), p(r), p(F), p(a), s.n, ctypes.byref(ph), sf.item(),
                 p(gx), p(sums), p(ws), ws.numel(), None)
    torch.cuda.synchronize()
    S = sums.cpu().numpy()
    assert _field_err(gx.cpu().numpy(), x.grad.cpu().numpy()) <= 1e-5
    # the parameter partials are sums of per-droplet terms of both signs
    # (random adjoint): errors are measured against sum |term|
    with torch.no_grad():
        _, F64, V64 = r64.velocity(x, r, m, eta, sf, s.L, return_parts=True)
        aF = (a * F64).sum(1).abs()
        scale_m = float(((a * (F64 / (sf * r[:, None]))).sum(1).abs().sum()
                         + (a * V64).sum(1).abs().sum()) / m)
        scale_eta = float((a * V64).sum(1).abs().sum() / eta)
        scale_sf = float((aF / r).sum() / sf**2)
    assert abs(S[0] / m.item() - m.grad.item()) <= 1e-5 * scale_m
    assert abs(S[1] / eta.item() - eta.grad.item()) <= 1e-5 * scale_eta
    assert abs(-S[2] / sf.item() - sf.grad.item()) <= 1e-5 * scale_sf
    # and relative to the value itself, at the trajectory tolerance
    assert _rel(S[0] / m.item(), m.grad.item()) <= 1e-4
    assert _rel(S[1] / eta.item(), eta.grad.item()) <= 1e-4
    assert _rel(-S[2] / sf.item(), sf.grad.item()) <= 1e-4


def test_contact_loss_matches_fp64(gpu):
    s = make_system(900, seed=4, phi=0.05)
    dev = torch.device("cuda")
    # pull a few droplets close to their neighbours so softplus terms matter
    pos = s.positions.copy()
    pos[1::7] = pos[0::7][: len(pos[1::7])] + 1.1 * (s.radii[0::7][: len(pos[1::7]), None]
                                                        + s.radii[1::7][:, None]) * np.array([0, 0, 1.0])
    pos %= s.L
    x = torch.tensor(pos, device=dev, requires_grad=True)
    r = torch.tensor(s.radii, device=dev)
    ref = r64.contact_loss(x, r, s.L)
    ref.backward()
    traj = DirectTrajectory(s.radii, s.L, 1e-3, 1)
    xg = torch.tensor(pos, device=dev, requires_grad=True)
    got = contact_loss(xg, traj)
    got.backward()
    assert ref.item() > 10
    assert _rel(got.item(), ref.item()) <= 1e-5
    assert _field_err(xg.grad.cpu().numpy(), x.grad.cpu().numpy()) <= 1e-5


def _dt(s, E, mu):
    fc = DirectDropletForceCalculator(num_particles=s.n, eps_oil=EPS, eta_oil=mu,
                                      eta_water=MU_W, E=E, L=s.L)
    F, V = fc.calculate_forces_and_convection(s.positions, s.radii)
    v = np.linalg.norm(F / (r64.stokes_factor(mu, MU_W) * s.radii[:, None]) + V, axis=1)
    return 0.1 * s.radii.min() / v.max()


def _fp64_loss(x0, r, E, mu, dt, T, L):
    xT = r64.rollout(x0, r, E, mu, EPS, MU_W, dt, T, L, checkpoint_every=10)
    return r64.contact_loss(xT, r, L)


def test_trajectory_gradients_match_autograd_and_fd(gpu):
    s = make_system(512, seed=0, phi=0.03)
    dev = torch.device("cuda")
    E0, mu0, T = 3e5, 0.05, 40
    dt = _dt(s, E0, mu0)
    traj = DirectTrajectory(s.radii, s.L, dt, T, eps_oil=EPS, eta_water=MU_W,
                            checkpoint_every=20)
    x0 = torch.tensor(s.positions, device=dev, requires_grad=True)
    E = torch.tensor(E0, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(mu0, dtype=torch.float64, requires_grad=True)
    loss = contact_loss(rollout(x0, E, mu, traj), traj)
    loss.backward()
    assert traj.recomputed_steps == T  # every segment recomputed once from its checkpoint

    r = torch.tensor(s.radii, device=dev)
    x64 = torch.tensor(s.positions, device=dev, requires_grad=True)
    E64 = torch.tensor(E0, dtype=torch.float64, device=dev, requires_grad=True)
    mu64 = torch.tensor(mu0, dtype=torch.float64, device=dev, requires_grad=True)
    L64 = _fp64_loss(x64, r, E64, mu64, dt, T, s.L)
    L64.backward()
    assert _rel(loss.item(), L64.item()) <= 1e-4
    assert _rel(E.grad.item(), E64.grad.item()) <= 1e-4
    assert _rel(mu.grad.item(), mu64.grad.item()) <= 1e-4
    assert _field_err(x0.grad.cpu().numpy(), x64.grad.cpu().numpy()) <= 1e-4

    # FP64 central differences: h = 1e-6 relative (positions h L), or the
    # largest smaller h whose two perturbed trajectories switch no contact
    # mask or periodic image (the loss jumps there)
    x0d = torch.tensor(s.positions, device=dev)
    g = torch.Generator(dev).manual_seed(3)
    v = torch.randn(s.n, 3, dtype=torch.float64, device=dev, generator=g)
    args = (x0d, r, E0, mu0, EPS, MU_W, dt, T, s.L)
    fE, hE = r64.switch_free_derivative(*args, "E")
    fmu, hmu = r64.switch_free_derivative(*args, "mu")
    fx, hx = r64.switch_free_derivative(*args, "x0", direction=v)
    assert None not in (fE, fmu, fx), "every step straddles a contact or image switch"
    assert hE == 1e-6 and hmu == 1e-6
    assert _rel(E.grad.item(), fE) <= 1e-4
    assert _rel(mu.grad.item(), fmu) <= 1e-4
    assert _rel(float((x0.grad * v).sum()), fx) <= 1e-4


def test_full_workload_gradients_match_autograd_and_fd(gpu):
    """The differentiable workload itself (N = 4096, T = 200, checkpoints
    every 20 steps): gradients w.r.t. E, mu_oil and x0 within 1e-4 of
    autograd through the plain FP64 path and of FP64 central differences."""
    s = differentiable()
    dev = torch.device("cuda")
    E0, mu0, T = 3e5, 0.05, 200
    dt = differentiable_dt(s, E0, mu0, EPS, MU_W)
    traj = DirectTrajectory(s.radii, s.L, dt, T, eps_oil=EPS, eta_water=MU_W,
                            checkpoint_every=20)
    x0 = torch.tensor(s.positions, device=dev, requires_grad=True)
    E = torch.tensor(E0, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(mu0, dtype=torch.float64, requires_grad=True)
    loss = contact_loss(rollout(x0, E, mu, traj), traj)
    loss.backward()
    r = torch.tensor(s.radii, device=dev)
    x64 = torch.tensor(s.positions, device=dev, requires_grad=True)
    E64 = torch.tensor(E0, dtype=torch.float64, device=dev, requires_grad=True)
    mu64 = torch.tensor(mu0, dtype=torch.float64, device=dev, requires_grad=True)
    L64 = _fp64_loss(x64, r, E64, mu64, dt, T, s.L)
    L64.backward()
    errs = {"loss": _rel(loss.item(), L64.item()),
            "E": _rel(E.grad.item(), E64.grad.item()),
            "mu": _rel(mu.grad.item(), mu64.grad.item()),
            "x0_field": _field_err(x0.grad.cpu().numpy(), x64.grad.cpu().numpy())}
    x0d = torch.tensor(s.positions, device=dev)
    g = torch.Generator(dev).manual_seed(3)
    v = torch.randn(s.n, 3, dtype=torch.float64, device=dev, generator=g)
    args = (x0d, r, E0, mu0, EPS, MU_W, dt, T, s.L)
    fE, hE = r64.switch_free_derivative(*args, "E")
    fmu, hmu = r64.switch_free_derivative(*args, "mu")
    fx, hx = r64.switch_free_derivative(*args, "x0", direction=v)
    assert None not in (fE, fmu, fx), "every step straddles a contact or image switch"
    errs.update({"fd_E": _rel(E.grad.item(), fE), "fd_mu": _rel(mu.grad.item(), fmu),
                 "fd_x0_dir": _rel(float((x0.grad * v).sum()), fx)})
    print(f"\n[adjoint N={s.n} T={T}] h E {hE:g}, mu {hmu:g}, x0 {hx:g}; errors {errs}")
    assert max(errs.values()) <= 1e-4, errs
