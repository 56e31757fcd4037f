"""torch.autograd wrapper of the buckling analysis: KS aggregate of the
first buckling loads as a differentiable function of the per-element
thicknesses and the nodal coordinates.

The forward pass runs the whole analysis on the GPU (static path,
shift-invert buckling, eigen-sensitivities through one path-adjoint solve)
and keeps the exact gradients; the backward pass scales them by the
incoming gradient.  Torch is only the calling convention here: tensors are
moved to host float64 arrays for the C ABI, and gradients come back on the
input tensors' devices."""
import numpy as np
import torch

from . import Context


class BucklingProblem:
    """Mesh, material, BCs and load of a linear buckling analysis."""

    def __init__(self, conn, Xpts, bcs, force, E=70e9, nu=0.3, nev=5, max_lanczos=100, tol=1e-12, rho=100.0,
                 sigma=0.0, ordering_seed=None):
        self.conn = np.ascontiguousarray(conn, np.int32)
        self.Xpts = np.ascontiguousarray(Xpts, np.float64)
        self.bcs, self.force = bcs, np.asarray(force, np.float64)
        self.E, self.nu = E, nu
        self.nev, self.max_lanczos, self.tol, self.rho, self.sigma = nev, max_lanczos, tol, rho, sigma
        self.ordering_seed = ordering_seed

    def evaluate(self, thick, Xpts=None, coords=False):
        """KS of -lambda / lambda_1 per design, its gradients w.r.t. the
        thicknesses [nb, ne] and (if coords) the coordinates [nb, 3 nn], and
        the loads [nb, nev]."""
        thick = np.atleast_2d(np.asarray(thick, np.float64))
        X = self.Xpts if Xpts is None else np.ascontiguousarray(Xpts, np.float64).reshape(-1, 3)
        ctx = Context(0)
        try:
            ctx.set_mesh(self.conn, X)
            ctx.set_material(self.E, self.nu)
            ctx.set_bcs(*self.bcs)
            ctx.set_designs(thick)
            if self.ordering_seed is not None:
                ctx.set_ordering_seed(self.ordering_seed)
            B = ctx.buckling(sigma=self.sigma, nev=self.nev, max_lanczos=self.max_lanczos, tol=self.tol)
            B.solve(force=self.force)
            lam = B.eigenvalues()[0]
            ks = np.zeros(lam.shape[0])
            coef = np.zeros_like(lam)
            for d in range(lam.shape[0]):
                ks[d], coef[d] = ks_of_loads(lam[d], self.rho)
            gt, gX = B.eigen_sens(coef, thickness=True, coords=coords)
            B.destroy()
        finally:
            ctx.close()
        return ks, gt, gX, lam


def ks_of_loads(lam, rho):
    """KS(rho) of f_i = -lambda_i / lambda_1 (a smooth minimum of the loads
    relative to the first) and d KS / d lambda_j, with the normalisation
    differentiated too: d f_i / d lambda_j = -delta_ij / lambda_1 +
    delta_1j lambda_i / lambda_1^2."""
    from . import ks_aggregate
    lam = np.asarray(lam, np.float64)
    ks, dks = ks_aggregate(-lam / lam[0], rho)
    coef = -dks / lam[0]
    coef[0] += float(dks @ lam) / lam[0] ** 2
    return ks, coef


class _BucklingKS(torch.autograd.Function):
    @staticmethod
    def forward(fctx, thick, Xpts, problem):
        coords = Xpts is not None and Xpts.requires_grad
        ks, gt, gX, lam = problem.evaluate(thick.detach().cpu().numpy(),
                                           None if Xpts is None else Xpts.detach().cpu().numpy(), coords)
        fctx.save_for_backward(torch.from_numpy(gt), torch.from_numpy(gX) if gX is not None else torch.zeros(0))
        fctx.coords = coords
        fctx.devices = (thick.device, Xpts.device if Xpts is not None else None)
        fctx.xshape = None if Xpts is None else Xpts.shape
        return torch.from_numpy(ks).to(thick.device)

    @staticmethod
    def backward(fctx, grad_ks):
        gt, gX = fctx.saved_tensors
        g = grad_ks.detach().cpu().to(torch.float64)
        dthick = (g[:, None] * gt).to(fctx.devices[0])
        dX = None
        if fctx.coords:
            # one coordinate set shared by all designs: sum over designs
            dX = (g[:, None] * gX).sum(0).reshape(fctx.xshape).to(fctx.devices[1])
        return dthick, dX, None


def buckling_ks(problem, thick, Xpts=None):
    """KS(rho) of -lambda_i / lambda_1 over the first nev buckling loads of
    every design (rows of thick [nb, ne]); differentiable in thick and, when
    Xpts [nn, 3] requires grad, in the coordinates."""
    if thick.dim() == 1:
        thick = thick[None, :]
    return _BucklingKS.apply(thick, Xpts, problem)
