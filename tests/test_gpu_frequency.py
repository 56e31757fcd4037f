"""Mass matrix and natural-frequency analysis (TACSFrequencyAnalysis, Lanczos
path) on the GPU against the oracle: the assembled mass matrix within 1e-12
(row-floor relative metric), the first 5 eigenvalues omega^2 within 1e-8 of a
dense generalised eigensolve of the oracle's K and M, M-normalised
eigenvectors with small residuals, the thickness and nodal-coordinate
gradients of a weighted sum of eigenvalues against central differences of the
dense eigenvalues, a shift above the first eigenvalue, and batched designs
against single designs."""
import numpy as np
import pytest
import scipy.linalg as sla

import pyoracle as O
import shellgpu as sg
from helpers import bsr_row_floor_ratio, random_mesh
from shellgpu.problems import Cylinder

pytestmark = pytest.mark.gpu

RHO = 2700.0


def make_ctx(conn, X, thick, bcs, seed_nodes=None):
    ctx = sg.Context(0)
    ctx.set_mesh(conn, X)
    ctx.set_material(70e9, 0.3, RHO, 0.0, 10.0)
    ctx.set_bcs(*bcs)
    ctx.set_designs(thick)
    if seed_nodes is not None:
        ctx.set_ordering_seed(seed_nodes)
    return ctx


def dense_frequencies(conn, X, thick, bcs, rowp, cols, nev):
    """Smallest nev eigenvalues of K x = lambda M x on the free dofs (dense):
    M is only semi-definite (the drilling rotation carries no mass), so the
    pencil is solved as M x = mu K x with lambda = 1 / mu."""
    m = O.material(rho=RHO)
    K = O.bsr_to_dense(rowp, cols, O.assemble_mat(O.KIND_K, 0, m, X, conn, thick, rowp, cols))
    M = O.bsr_to_dense(rowp, cols, O.assemble_mat(O.KIND_M, 0, m, X, conn, thick, rowp, cols))
    fixed = np.zeros(K.shape[0], bool)
    for node, mask in zip(*bcs):
        for k in range(6):
            if mask & (1 << k):
                fixed[6 * node + k] = True
    free = ~fixed
    mu, V = sla.eigh(M[np.ix_(free, free)], K[np.ix_(free, free)])
    order = np.argsort(-mu)[:nev]
    return 1.0 / mu[order]


def test_mass_matrix_matches_oracle():
    for conn, X, bcs in [(Cylinder(6, 10).conn, Cylinder(6, 10).Xpts, Cylinder(6, 10).bcs()),
                         random_mesh(5, 4, 2) + ((np.array([0, 1], np.int32), np.array([0x3F, 0x3F], np.int32)),)]:
        thick = np.random.default_rng(3).uniform(0.0015, 0.0025, conn.shape[0])
        ctx = make_ctx(conn, X, thick, bcs)
        rowp, cols = ctx.pattern()
        A = ctx.matrix()
        ctx.assemble_mattype(sg.MASS, A)
        ref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_M, 0, O.material(rho=RHO), X, conn, thick, rowp, cols),
                              rowp, cols, *bcs)
        assert bsr_row_floor_ratio(rowp, A.values()[0], ref) < 1e-12


@pytest.mark.parametrize("sigma_factor,block", [(0.0, False), (0.8, False), (3.0, False), (0.0, True)])
def test_frequencies_match_dense_oracle(sigma_factor, block, monkeypatch):
    if block:  # the sequential factorisation with block Lanczos (blocks of 4)
        monkeypatch.setenv("SHELLGPU_FACTOR", "thomas")
        monkeypatch.setenv("SHELLGPU_LANCZOS_BLOCK", "4")
    cyl = Cylinder(10, 20)
    thick = np.random.default_rng(4).uniform(0.0015, 0.0025, cyl.nelems)
    ctx = make_ctx(cyl.conn, cyl.Xpts, thick, cyl.bcs(), cyl.ring_nodes(0))
    rowp, cols = ctx.pattern()
    nev = 8
    ref = dense_frequencies(cyl.conn, cyl.Xpts, thick, cyl.bcs(), rowp, cols, nev)
    # sigma_factor 3: the shift lies above the first eigenvalue, so the
    # shifted operator is indefinite and the shift is lowered until it factors
    fr = ctx.frequency(sigma=sigma_factor * ref[0], nev=nev, max_lanczos=150, tol=1e-12)
    fr.solve()
    lam, err = fr.eigenvalues()
    info = fr.info()
    print("gpu", lam[0][:5], "dense", ref[:5], "steps", info["steps"], "sigma", info["sigma"])
    assert np.abs(lam[0][:5] - ref[:5]).max() <= 1e-8 * np.abs(ref[:5]).max()
    assert err[0][:5].max() <= 1e-10
    if sigma_factor > 1.0:
        assert info["sigma"][0] < ref[0]
    # eigenvectors: M-normalised and K phi = lambda M phi
    m = O.material(rho=RHO)
    Kv = O.bsr_apply_bcs(O.assemble_mat(O.KIND_K, 0, m, cyl.Xpts, cyl.conn, thick, rowp, cols),
                         rowp, cols, *cyl.bcs())
    Mv = O.bsr_apply_bcs(O.assemble_mat(O.KIND_M, 0, m, cyl.Xpts, cyl.conn, thick, rowp, cols),
                         rowp, cols, *cyl.bcs())
    for j in range(5):
        phi = fr.eigenvector(j)[0]
        Kp, Mp = O.bsr_matvec(rowp, cols, Kv, phi), O.bsr_matvec(rowp, cols, Mv, phi)
        assert abs(phi @ Mp - 1.0) <= 1e-10
        assert np.linalg.norm(Kp - lam[0][j] * Mp) <= 1e-7 * np.linalg.norm(Kp)
    fr.destroy()
    ctx.close()


def test_frequency_thickness_gradient_matches_differences():
    cyl = Cylinder(8, 16)
    rng = np.random.default_rng(5)
    thick = rng.uniform(0.0015, 0.0025, cyl.nelems)
    ctx = make_ctx(cyl.conn, cyl.Xpts, thick, cyl.bcs(), cyl.ring_nodes(0))
    rowp, cols = ctx.pattern()
    nev = 6
    fr = ctx.frequency(sigma=0.0, nev=nev, max_lanczos=150, tol=1e-12)
    fr.solve()
    lam, _ = fr.eigenvalues()
    coef = np.array([1.0, 0.5, 0.25, 0.0, 0.0, 0.0])
    grad = fr.eigen_sens(coef)[0]
    ref = dense_frequencies(cyl.conn, cyl.Xpts, thick, cyl.bcs(), rowp, cols, nev)
    gaps = np.diff(ref[:4]) / ref[:3]
    print("eigenvalues", ref[:5], "relative gaps", gaps)
    assert gaps.min() > 1e-4  # the random thickness separates the cylinder's pairs
    for k in range(3):
        p = rng.standard_normal(cyl.nelems)
        h = 1e-4 * 0.002

        def f(s):
            return coef[:3] @ dense_frequencies(cyl.conn, cyl.Xpts, thick + s * p, cyl.bcs(), rowp, cols, 3)

        fd1 = (f(h) - f(-h)) / (2 * h)
        fd2 = (f(2 * h) - f(-2 * h)) / (4 * h)
        fd = (4 * fd1 - fd2) / 3  # Richardson
        an = grad @ p
        print("directional", an, fd)
        assert abs(an - fd) <= 1e-6 * abs(fd)
    fr.destroy()
    ctx.close()


def test_frequency_coordinate_gradient_matches_differences():
    # d(lambda)/dX = phi^T (dK/dX - lambda dM/dX) phi / (phi^T M phi), the
    # formula of TACSFrequencyAnalysis::evalEigenXptSens; checked along random
    # nodal-coordinate directions against Richardson-extrapolated central
    # differences of the dense eigenvalues of the moved mesh
    cyl = Cylinder(8, 16)
    rng = np.random.default_rng(7)
    thick = rng.uniform(0.0015, 0.0025, cyl.nelems)
    ctx = make_ctx(cyl.conn, cyl.Xpts, thick, cyl.bcs(), cyl.ring_nodes(0))
    rowp, cols = ctx.pattern()
    nev = 6
    fr = ctx.frequency(sigma=0.0, nev=nev, max_lanczos=150, tol=1e-12)
    fr.solve()
    coef = np.array([1.0, 0.5, 0.25, 0.0, 0.0, 0.0])
    gt, gX = fr.eigen_sens(coef, xpts=True)
    # the thickness part of the same call is the thickness-only result
    assert np.abs(gt[0] - fr.eigen_sens(coef)[0]).max() <= 1e-12 * np.abs(gt[0]).max()
    ref = dense_frequencies(cyl.conn, cyl.Xpts, thick, cyl.bcs(), rowp, cols, nev)
    assert (np.diff(ref[:4]) / ref[:3]).min() > 1e-4
    for k in range(3):
        p = rng.standard_normal(cyl.Xpts.shape)
        h = 1e-5

        def f(s):
            return coef[:3] @ dense_frequencies(cyl.conn, cyl.Xpts + s * p, thick, cyl.bcs(), rowp, cols, 3)

        fd1 = (f(h) - f(-h)) / (2 * h)
        fd2 = (f(2 * h) - f(-2 * h)) / (4 * h)
        fd = (4 * fd1 - fd2) / 3
        an = np.sum(gX[0] * p)
        print("directional", an, fd)
        assert abs(an - fd) <= 1e-6 * abs(fd)
    fr.destroy()
    # moving the nodes of the same context (TACSAssembler::setNodes): the
    # eigenvalues of the moved mesh, and its coordinates read back exactly
    Xm = cyl.Xpts + 1e-3 * rng.standard_normal(cyl.Xpts.shape)
    ctx.set_nodes(Xm)
    assert np.array_equal(ctx.get_nodes(), Xm)
    fr = ctx.frequency(sigma=0.0, nev=nev, max_lanczos=150, tol=1e-12)
    fr.solve()
    lam, _ = fr.eigenvalues()
    refm = dense_frequencies(cyl.conn, Xm, thick, cyl.bcs(), rowp, cols, nev)
    assert np.abs(lam[0][:5] - refm[:5]).max() <= 1e-8 * np.abs(refm[:5]).max()
    assert np.abs(refm[0] - ref[0]) > 1e-6 * ref[0]  # the move changed the answer
    fr.destroy()
    ctx.close()


def test_batched_designs_match_single_designs():
    cyl = Cylinder(8, 16)
    rng = np.random.default_rng(6)
    thick = rng.uniform(0.0015, 0.0025, (3, cyl.nelems))
    ctx = make_ctx(cyl.conn, cyl.Xpts, thick, cyl.bcs(), cyl.ring_nodes(0))
    fr = ctx.frequency(sigma=0.0, nev=5, max_lanczos=150, tol=1e-12)
    fr.solve()
    lam, _ = fr.eigenvalues()
    g = fr.eigen_sens(np.ones(5))
    fr.destroy()
    ctx.close()
    for d in range(3):
        c1 = make_ctx(cyl.conn, cyl.Xpts, thick[d], cyl.bcs(), cyl.ring_nodes(0))
        f1 = c1.frequency(sigma=0.0, nev=5, max_lanczos=150, tol=1e-12)
        f1.solve()
        l1, _ = f1.eigenvalues()
        g1 = f1.eigen_sens(np.ones(5))
        assert np.abs(lam[d] - l1[0]).max() <= 1e-10 * np.abs(l1[0]).max()
        assert np.abs(g[d] - g1[0]).max() <= 1e-8 * np.abs(g1[0]).max()
        f1.destroy()
        c1.close()
