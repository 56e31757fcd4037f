"""torch.autograd harness of the buckling KS: backward gradients along
seeded directions against central differences of the forward pass, for
thickness and coordinates."""
import numpy as np
import pytest
import torch

from shellgpu.autograd import BucklingProblem, buckling_ks
from shellgpu.problems import Cylinder, seeded_thickness

pytestmark = pytest.mark.gpu


def test_autograd_ks_gradients_match_differences():
    cyl = Cylinder(8, 16)
    prob = BucklingProblem(cyl.conn, cyl.Xpts, cyl.bcs(), cyl.edge_load(1.0), nev=4, max_lanczos=80,
                           ordering_seed=cyl.ring_nodes(0))
    t0 = seeded_thickness(cyl.nelems, 81, ndesigns=2)
    thick = torch.tensor(t0, dtype=torch.float64, requires_grad=True)
    X = torch.tensor(cyl.Xpts, dtype=torch.float64, requires_grad=True)
    ks = buckling_ks(prob, thick, X)
    w = torch.tensor([1.0, -0.5], dtype=torch.float64)
    (w * ks).sum().backward()
    rng = np.random.default_rng(3003)
    dt = rng.standard_normal(t0.shape) * t0
    dX = rng.standard_normal(cyl.Xpts.shape) * 1e-3

    def f(tv, Xv):
        return float((w.numpy() * prob.evaluate(tv, Xv)[0]).sum())

    # the eigenvalues carry round-off of ~1e-13 relative, which puts ~1e-8 of
    # noise on a central difference at h = 1e-5; h = 1e-4 with Richardson
    # extrapolation keeps both the noise and the truncation error below
    # 1e-9 of the derivative
    h = 1e-4

    def cd(hh):
        return (f(t0 + hh * dt, cyl.Xpts) - f(t0 - hh * dt, cyl.Xpts)) / (2 * hh)

    fd_t = (4.0 * cd(h) - cd(2 * h)) / 3.0
    ad_t = float((thick.grad.numpy() * dt).sum())
    print("thickness", ad_t, fd_t)
    assert abs(ad_t - fd_t) <= 1e-5 * max(abs(fd_t), 1e-3)
    h = 1e-4
    fd_X = (f(t0, cyl.Xpts + h * dX) - f(t0, cyl.Xpts - h * dX)) / (2 * h)
    ad_X = float((X.grad.numpy() * dX).sum())
    print("coords", ad_X, fd_X)
    assert abs(ad_X - fd_X) <= 1e-5 * max(abs(fd_X), 1e-3)
