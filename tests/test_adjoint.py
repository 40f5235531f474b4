"""Adjoint of the RHS and RK4 step (axs_adjoint_grad_energy) against central
finite differences of the forward objective and against the FP64 adjoint,
through the torch.autograd binding."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def problem(ctx, n=16, precision=0, nsteps=8, K=3):
    from axion_strings_adaptive_amd import capi
    from axion_strings_adaptive_amd.torch_adjoint import GradEnergyProblem
    bg = capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.0,
                              has_switch=True)
    return GradEnergyProblem(ctx, n, 0.5, precision, bg, 1.0, 0.1, nsteps, K)


def test_adjoint_matches_central_differences_fp64(ctx, oracle):
    n = 16
    p = problem(ctx, n)
    s0 = oracle.seeded_ic(n, seed=5)
    s0[2:] = 0.1 * np.random.default_rng(1).normal(size=s0[2:].shape)
    J, g, gl = p.value_and_grad(s0, 1.0)
    assert J == pytest.approx(p.forward_value(s0, 1.0), rel=1e-14)
    rng = np.random.default_rng(2)
    eps = 1e-3
    for _ in range(8):
        v = rng.normal(size=s0.shape)
        fd = (p.forward_value(s0 + eps * v, 1.0) -
              p.forward_value(s0 - eps * v, 1.0)) / (2 * eps)
        assert np.sum(g * v) == pytest.approx(fd, rel=1e-5, abs=1e-9 * abs(J))
    fdl = (p.forward_value(s0, 1.0 + eps) - p.forward_value(s0, 1.0 - eps)) / (
        2 * eps)
    assert gl == pytest.approx(fdl, rel=1e-5)


def test_checkpoint_interval_does_not_change_the_gradient(ctx, oracle):
    n = 16
    s0 = oracle.seeded_ic(n, seed=6)
    J1, g1, l1 = problem(ctx, n, K=1).value_and_grad(s0, 1.0)
    J2, g2, l2 = problem(ctx, n, K=5).value_and_grad(s0, 1.0)
    J3, g3, l3 = problem(ctx, n, K=8).value_and_grad(s0, 1.0)
    assert J1 == J2 == J3
    assert np.array_equal(g1, g2) and np.array_equal(g1, g3)
    assert l1 == l2 == l3


def test_fp32_forward_gradient_tracks_fp64(ctx, oracle):
    from axion_strings_adaptive_amd import capi
    n = 16
    s0 = oracle.seeded_ic(n, seed=7)
    J64, g64, l64 = problem(ctx, n, precision=capi.AXS_FP64).value_and_grad(
        s0, 1.0)
    J32, g32, l32 = problem(ctx, n, precision=capi.AXS_FP32C).value_and_grad(
        s0, 1.0)
    assert J32 == pytest.approx(J64, rel=1e-5)
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-4
    assert l32 == pytest.approx(l64, rel=1e-3)


def test_torch_autograd_binding(ctx, oracle):
    import torch
    from axion_strings_adaptive_amd.torch_adjoint import GradientEnergy
    n = 16
    p = problem(ctx, n, nsteps=4, K=2)
    s0 = torch.tensor(oracle.seeded_ic(n, seed=8), dtype=torch.float64,
                      requires_grad=True)
    lam = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    J = GradientEnergy.apply(s0, lam, p)
    (3.0 * J).backward()
    _, g, gl = p.value_and_grad(s0.detach().numpy(), 1.0)
    np.testing.assert_allclose(s0.grad.numpy(), 3.0 * g, rtol=1e-14)
    assert lam.grad.item() == pytest.approx(3.0 * gl, rel=1e-14)
