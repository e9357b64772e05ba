"""Host-side pieces of the adjoint path (no GPU): the differentiable two-stream
initial state and the checkpoint schedule arithmetic."""

import math

import numpy as np
import torch

from pypic3d_amd import workloads
from pypic3d_amd.adjoint import _inverse_cdf, two_stream_state


def test_inverse_cdf_solves_the_perturbed_cdf():
    W, amp = 12.0, 0.2
    u = torch.linspace(0.0, W * (1 - 1e-9), 101, dtype=torch.float64)
    y = _inverse_cdf(u, W, torch.tensor(amp, dtype=torch.float64), 1)
    k = 2 * math.pi / W
    np.testing.assert_allclose((y + (amp / k) * torch.sin(k * y)).numpy(), u.numpy(), atol=1e-12)
    assert torch.all(y[1:] > y[:-1])


def test_two_stream_state_derivatives_match_central_differences():
    static, dyn = workloads.make_parameters((12, 10, 1), cfl=0.5)
    d = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(0.05, dtype=torch.float64, requires_grad=True)
    x0, v0 = two_stream_state(static, dyn, 2, d, a, seed=3, device="cpu")
    assert x0.shape == (3, 2, 240) and v0.shape == (3, 2, 240)
    w = torch.randn(x0.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(0))
    f = (w * x0).sum() + (w * v0).sum()
    gd, ga = torch.autograd.grad(f, [d, a])
    h = 1e-6

    def F(dr, am):
        x, v = two_stream_state(static, dyn, 2, dr, am, seed=3, device="cpu")
        return float((w * x).sum() + (w * v).sum())

    assert abs(float(gd) - (F(0.1 + h, 0.05) - F(0.1 - h, 0.05)) / (2 * h)) < 1e-6
    assert abs(float(ga) - (F(0.1, 0.05 + h) - F(0.1, 0.05 - h)) / (2 * h)) < 1e-6
    # beams drift in opposite directions; the reduced axis stays inside its cell
    assert abs(float(v0[0, 0].detach().mean()) - 0.1) < 0.01 and abs(float(v0[0, 1].detach().mean()) + 0.1) < 0.01
    assert float(x0[2].detach().abs().max()) < 0.5
