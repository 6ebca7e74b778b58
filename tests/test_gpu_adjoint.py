"""Adjoint gradients against FP64 central finite differences (h = 1e-6) of the same
forward map: 512x512 BGK, tau = 0.6, 500 steps, checkpoints every 50, band-limited
initial velocity (|k| <= 8, amplitude 0.02, seed 1), target from seed 2.

FP64 adjoint: <grad, d> within 1e-6 relative of the FD quotient on 16 random directions of
the initial field, and for tau (measured on MI355X: 1.6e-9 and 2.7e-11).  FP32 forward and
checkpoints with the FP64 adjoint state: within 6e-6, twice the measured 2.9e-6 (tau:
1.5e-6); that is the deviation of the FP32 trajectory's gradient from the FP64 one."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, STEPS, EVERY, TAU, H = 512, 500, 50, 0.6, 1e-6


def setup(P, torch, storage, n=N, steps=STEPS):
    ux0, uy0 = P.band_limited_field(n, n, 8, 0.02, 1)
    uxt, uyt = P.band_limited_field(n, n, 8, 0.02, 2)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(np.ravel(a, order="F"))).cuda().view(n, n)
    eng = P.Adjoint(n, n, steps, EVERY, storage)
    return eng, t(ux0), t(uy0), t(uxt), t(uyt)


def directions(torch, n, k, seed=123):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return [(torch.randn(n, n, generator=g, dtype=torch.float64).cuda(),
             torch.randn(n, n, generator=g, dtype=torch.float64).cuda()) for _ in range(k)]


def fd_loss(eng64, rho0, ux0, uy0, tau, uxt, uyt):
    return eng64.forward(rho0, ux0.contiguous(), uy0.contiguous(), 1.0 / tau, uxt, uyt)


@pytest.mark.parametrize("storage,tol", [("f64", 1e-6), ("f32s", 6e-6)])
def test_adjoint_matches_central_fd(P, gpu, storage, tol):
    import torch
    print(f"case adjoint {storage} {N}^2 x {STEPS}", flush=True)
    eng, ux0, uy0, uxt, uyt = setup(P, torch, storage)
    eng64, *_ = setup(P, torch, "f64")
    rho0 = torch.ones_like(ux0)
    loss = eng.forward(rho0, ux0, uy0, 1.0 / TAU, uxt, uyt)
    _, gux, guy, gom = eng.backward()
    gtau = -gom / TAU ** 2
    assert np.isfinite(loss) and loss > 0
    worst = 0.0
    for dx, dy in directions(torch, N, 16):
        lp = fd_loss(eng64, rho0, ux0 + H * dx, uy0 + H * dy, TAU, uxt, uyt)
        lm = fd_loss(eng64, rho0, ux0 - H * dx, uy0 - H * dy, TAU, uxt, uyt)
        fd = (lp - lm) / (2 * H)
        ad = float(torch.sum(gux * dx) + torch.sum(guy * dy))
        worst = max(worst, abs(ad - fd) / abs(fd))
    lp = fd_loss(eng64, rho0, ux0, uy0, TAU + H, uxt, uyt)
    lm = fd_loss(eng64, rho0, ux0, uy0, TAU - H, uxt, uyt)
    fd_tau = (lp - lm) / (2 * H)
    err_tau = abs(gtau - fd_tau) / abs(fd_tau)
    print(f"{storage}: worst field-direction rel err {worst:.2e}, tau rel err {err_tau:.2e}")
    assert worst < tol
    assert err_tau < tol


def test_autograd_function(P, gpu):
    import torch
    n, steps = 64, 60
    eng, ux0, uy0, uxt, uyt = setup(P, torch, "f64", n=n, steps=steps)
    ux = ux0.clone().requires_grad_(True)
    uy = uy0.clone().requires_grad_(True)
    tau = torch.tensor(TAU, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = P.velocity_misfit(ux, uy, tau, uxt, uyt, eng)
    loss.backward()
    eng2, *_ = setup(P, torch, "f64", n=n, steps=steps)
    rho0 = torch.ones_like(ux0)
    l2 = eng2.forward(rho0, ux0, uy0, 1.0 / TAU, uxt, uyt)
    _, gux, guy, gom = eng2.backward()
    assert float(loss.detach()) == l2
    torch.testing.assert_close(ux.grad, gux, rtol=0, atol=0)
    torch.testing.assert_close(uy.grad, guy, rtol=0, atol=0)
    assert float(tau.grad) == pytest.approx(-gom / TAU ** 2, rel=1e-15)
    # one gradient-descent step on the initial field lowers the misfit
    with torch.no_grad():
        lr = 0.5 * float(loss.detach()) / float(torch.sum(ux.grad ** 2) + torch.sum(uy.grad ** 2))
        l_new = P.velocity_misfit(ux - lr * ux.grad, uy - lr * uy.grad, tau, uxt, uyt, eng)
    assert float(l_new) < float(loss.detach())


def test_checkpoint_interval_does_not_change_gradients(P, gpu):
    import torch
    n, steps = 48, 37
    grads = []
    for every in (1, 5, 37, 100):
        ux0, uy0 = P.band_limited_field(n, n, 4, 0.03, 5)
        uxt, uyt = P.band_limited_field(n, n, 4, 0.03, 6)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(np.ravel(a, order="F"))).cuda()
        eng = P.Adjoint(n, n, steps, every, "f64")
        rho0 = torch.ones(n * n, dtype=torch.float64, device="cuda")
        eng.forward(rho0, t(ux0), t(uy0), 1.7, t(uxt), t(uyt))
        grads.append(eng.backward())
    for g in grads[1:]:
        torch.testing.assert_close(g[1], grads[0][1], rtol=0, atol=0)
        assert g[3] == grads[0][3]
