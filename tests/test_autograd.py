"""The autograd bridge: backward contracts the upstream gradient with dT/dp
(checked on the CPU with a stand-in run), and on the GPU the gradient of a
scalar loss equals the contraction of the run's own dT/dp tallies."""
import importlib

import numpy as np
import pytest
import torch


class FakeModel:
    """linear T(p) = T0 + sum_s p_s G_s with known G"""

    def __init__(self):
        rng = np.random.default_rng(0)
        self.G = rng.normal(size=(4, 5, 3))

    def run(self, p):
        return 300.0 + np.tensordot(p, self.G, axes=1), self.G.copy()


def test_backward_contracts_dTdp(pkg):
    ag = importlib.import_module("bte_barna_amd.autograd")
    m = FakeModel()
    p = torch.tensor([0.0, 0.7, 0.3, 0.0], dtype=torch.float64, requires_grad=True)
    T = ag._SpecularityT.apply(p, m)
    w = torch.linspace(0, 1, 15, dtype=torch.float64).reshape(5, 3)
    (T * w).sum().backward()
    expect = (m.G * w.numpy()[None]).sum(axis=(1, 2))
    np.testing.assert_allclose(p.grad.numpy(), expect, rtol=1e-14)
    # torch's own check of the linear map
    assert torch.autograd.gradcheck(lambda q: ag._SpecularityT.apply(q, m), (p,), eps=1e-6, atol=1e-8)


@pytest.mark.gpu
def test_gpu_gradient_through_autograd(pkg, ctx):
    ag = importlib.import_module("bte_barna_amd.autograd")
    W = pkg.workloads
    mat = W.material(400, seed=2, mirror=True)
    rect, kind, teq, _ = W.ribbon(100.0, 20.0, 8, 300.0, reservoirs=(300.5, 299.5))
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    model = ag.SpecularityModel(pkg, ctx, modes, geom, 300.0, 50000, 1.0, 20.0, 200.0, seed=3, sigma_bounce=2.0)
    p = torch.tensor([0.0, 0.7, 0.3, 0.0], dtype=torch.float64, requires_grad=True)
    T = model(p)
    loss = T[-1, :4].mean()
    loss.backward()
    _, dTdp = model.run(p.detach().numpy())  # same seed: same tallies
    expect = dTdp[:, -1, :4].mean(axis=1)
    np.testing.assert_allclose(p.grad.numpy(), expect, rtol=1e-9, atol=1e-12)
    assert p.grad[0] == 0 and p.grad[3] == 0 and p.grad[1] != 0 and p.grad[2] != 0
