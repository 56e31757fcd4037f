"""Thickness and coordinate sensitivities on the GPU against the oracle.

Element level: phi^T (dK/dt) phi and phi^T (dG/dt) phi against a five-point
central difference of the oracle's element matrices in t (exact for the
cubic thickness dependence of the section: A ~ t, D ~ t^3), the path
gradient d(phi^T G phi)/du against differences of oracle G matrices (G is
affine in the path), and psi^T dR/dt against the oracle's restatement of
addAdjResProduct, both element models.

Coordinates: d/dX of phi^T K phi, phi^T G(u) phi, phi^T M phi, psi^T M phi
and psi^T R(u) (both models) summed over the mesh against five-point differences of the
oracle's element matrices and residuals in every nodal coordinate.

Analysis level: d(lambda_j)/dt and d(lambda_j)/dX of the linear buckling
loads and of a KS aggregate along seeded random directions against central
differences of the oracle's buckling loads, and batched designs against
one-design runs."""
import numpy as np
import pytest

import pyoracle as O
import shellgpu as sg
from helpers import random_mesh
from shellgpu.problems import Cylinder, seeded_thickness

pytestmark = pytest.mark.gpu


def _five_point(f, t, h):
    return (f(t - 2 * h) - 8 * f(t - h) + 8 * f(t + h) - f(t + 2 * h)) / (12 * h)


def _mesh_ctx(conn, X, thick, temp=None):
    ctx = sg.Context(0)
    ctx.set_mesh(conn, X)
    ctx.set_material(70e9, 0.3, 2700.0, 1.2e-5 if temp is not None else 0.0, 10.0)
    ctx.set_designs(thick, temp)
    return ctx


@pytest.mark.parametrize("thermal", [False, True])
def test_element_products_match_oracle(thermal):
    conn, X = random_mesh(5, 4, seed=11)
    nn, ne = X.shape[0], conn.shape[0]
    rng = np.random.default_rng(12)
    nb = 2
    thick = rng.uniform(0.002, 0.02, (nb, ne))
    temp = rng.uniform(-30.0, 60.0, (nb, ne)) if thermal else None
    mat = O.material(cte=1.2e-5 if thermal else 0.0)
    phi = rng.standard_normal((nb, 6 * nn))
    u = 1e-3 * rng.standard_normal((nb, 6 * nn))
    psi = rng.standard_normal((nb, 6 * nn))
    ctx = _mesh_ctx(conn, X, thick, temp)
    dphi, du, dpsi = ctx.vector(phi), ctx.vector(u), ctx.vector(psi)

    dK, dG = ctx.design_vector(), ctx.design_vector()
    ctx.add_mat_dv_sens_inner_product(dK, sg.STIFFNESS, dphi, scale=1.0)
    ctx.add_mat_dv_sens_inner_product(dG, sg.GEOMETRIC, dphi, du, scale=1.0)
    dRl, dRn = ctx.design_vector(), ctx.design_vector()
    ctx.add_adj_res_product(dRl, dpsi, du, model=sg.LINEAR)
    ctx.add_adj_res_product(dRn, dpsi, du, model=sg.NONLINEAR)
    gK, gG = dK.download().reshape(nb, ne), dG.download().reshape(nb, ne)
    gRl, gRn = dRl.download().reshape(nb, ne), dRn.download().reshape(nb, ne)

    rK, rG, rRl, rRn = (np.zeros((nb, ne)) for _ in range(4))
    for d in range(nb):
        for e in range(ne):
            Xe = X[conn[e]].ravel()
            idx = (6 * conn[e][:, None] + np.arange(6)).ravel()
            pe, ue, se = phi[d, idx], u[d, idx], psi[d, idx]
            T = temp[d, e] if thermal else 0.0
            t = thick[d, e]
            rK[d, e] = _five_point(lambda s: pe @ O.elem_stiffness(0, mat, s, T, Xe, np.zeros(24)) @ pe, t, 0.05 * t)
            rG[d, e] = _five_point(lambda s: pe @ O.elem_geometric_step(mat, s, T, Xe, ue, 1.0) @ pe, t, 0.05 * t)
            rRl[d, e] = O.elem_adj_res_dv(0, mat, t, T, se, Xe, ue)
            rRn[d, e] = O.elem_adj_res_dv(1, mat, t, T, se, Xe, ue)
    for name, g, r in (("K", gK, rK), ("G", gG, rG), ("R lin", gRl, rRl), ("R nl", gRn, rRn)):
        err = np.abs(g - r).max() / np.abs(r).max()
        print(name, err)
        assert err <= 1e-10, name


def test_path_gradient_matches_oracle():
    conn, X = random_mesh(4, 4, seed=21)
    nn, ne = X.shape[0], conn.shape[0]
    rng = np.random.default_rng(22)
    thick = rng.uniform(0.002, 0.02, (1, ne))
    temp = rng.uniform(-30.0, 60.0, (1, ne))
    mat = O.material(cte=1.2e-5)
    phi = rng.standard_normal(6 * nn)
    ctx = _mesh_ctx(conn, X, thick, temp)
    g = ctx.vector()
    ctx.eval_mat_sv_sens_inner_product(ctx.vector(phi), g)
    g = g.download()
    rowp, cols = O.node_csr(nn, conn)

    def pGp(u):
        G = O.assemble_mat(O.KIND_G, 0, mat, X, conn, thick[0], rowp, cols, u=u, temp=temp[0])
        return phi @ O.bsr_matvec(rowp, cols, G, phi)

    base = pGp(np.zeros(6 * nn))
    for k in range(4):
        w = rng.standard_normal(6 * nn)
        ref = pGp(w) - base  # G is affine in the path
        assert abs(g @ w - ref) <= 1e-10 * max(abs(ref), np.abs(g).max() * np.abs(w).sum() * 1e-6), k


@pytest.mark.parametrize("thermal", [False, True])
def test_coordinate_products_match_oracle(thermal):
    conn, X = random_mesh(3, 3, seed=51)
    nn, ne = X.shape[0], conn.shape[0]
    rng = np.random.default_rng(52)
    thick = rng.uniform(0.002, 0.02, (1, ne))
    temp = rng.uniform(-30.0, 60.0, (1, ne)) if thermal else None
    mat = O.material(cte=1.2e-5 if thermal else 0.0)
    phi = rng.standard_normal(6 * nn)
    u = 1e-3 * rng.standard_normal(6 * nn)
    psi = rng.standard_normal(6 * nn)
    ctx = _mesh_ctx(conn, X, thick, temp)
    dphi, du, dpsi = ctx.vector(phi), ctx.vector(u), ctx.vector(psi)
    out = {}
    for name in ("K", "G", "M", "Mpsi", "R0", "R1"):
        g = ctx.coord_vector()
        if name == "K":
            ctx.add_mat_xpt_sens_inner_product(g, sg.STIFFNESS, dphi)
        elif name == "M":
            ctx.add_mat_xpt_sens_inner_product(g, sg.MASS, dphi)
        elif name == "Mpsi":
            ctx.add_mat_xpt_sens_inner_product(g, sg.MASS, dphi, psi=dpsi)
        elif name == "G":
            ctx.add_mat_xpt_sens_inner_product(g, sg.GEOMETRIC, dphi, du)
        else:
            ctx.add_adj_res_xpt_product(g, dpsi, du, model=int(name[1]))
        out[name] = g.download()

    def total(name, Xp):
        s = 0.0
        for e in range(ne):
            Xe = Xp[conn[e]].ravel()
            idx = (6 * conn[e][:, None] + np.arange(6)).ravel()
            T = temp[0, e] if thermal else 0.0
            t = thick[0, e]
            if name == "K":
                s += phi[idx] @ O.elem_stiffness(0, mat, t, T, Xe, np.zeros(24)) @ phi[idx]
            elif name == "M":
                s += phi[idx] @ O.elem_mass(mat, t, Xe) @ phi[idx]
            elif name == "Mpsi":
                s += psi[idx] @ O.elem_mass(mat, t, Xe) @ phi[idx]
            elif name == "G":
                s += phi[idx] @ O.elem_geometric_step(mat, t, T, Xe, u[idx], 1.0) @ phi[idx]
            else:
                s += psi[idx] @ O.elem_residual(int(name[1]), mat, t, T, Xe, u[idx])
        return s

    h = 1e-3 * 0.3
    for name, g in out.items():
        ref = np.zeros(3 * nn)
        for k in range(3 * nn):
            def f(x, k=k):
                Xp = X.copy().ravel()
                Xp[k] = x
                return total(name, Xp.reshape(-1, 3))
            ref[k] = _five_point(f, X.ravel()[k], h)
        err = np.abs(g - ref).max() / np.abs(ref).max()
        print(name, err)
        assert err <= 1e-8, name


def _cyl_ctx(cyl, thick):
    ctx = sg.Context(0)
    ctx.set_mesh(cyl.conn, cyl.Xpts)
    ctx.set_material(70e9, 0.3)
    ctx.set_bcs(*cyl.bcs())
    ctx.set_designs(thick)
    ctx.set_ordering_seed(cyl.ring_nodes(0))
    return ctx


def test_buckling_gradient_matches_oracle_fd():
    cyl = Cylinder(10, 20)
    thick = seeded_thickness(cyl.nelems, 31)
    ctx = _cyl_ctx(cyl, thick)
    nev = 4
    B = ctx.buckling(sigma=0.0, nev=nev, max_lanczos=80, tol=1e-12)
    f = cyl.edge_load(1.0)
    B.solve(force=f)
    lam, err = B.eigenvalues()
    sig = B.info()["sigma"][0]
    grads = [B.eigen_sens(np.eye(nev)[j])[0][0] for j in range(3)]
    # KS aggregate of -lambda / sigma (tracks the smallest load), rho = 100
    ks, dks = sg.ks_aggregate(-lam[0] / sig, 100.0)
    gks = B.eigen_sens(-dks / sig)[0][0]

    rng = np.random.default_rng(3000)
    h = 1e-5
    for k in range(2):
        dirn = rng.standard_normal(cyl.nelems) * thick[0]
        lp = O.linear_buckling(cyl.Xpts, cyl.conn, thick[0] + h * dirn, cyl.bcs(), f, sig, nev=nev, max_lanczos=80)[0]
        lm = O.linear_buckling(cyl.Xpts, cyl.conn, thick[0] - h * dirn, cyl.bcs(), f, sig, nev=nev, max_lanczos=80)[0]
        fd = (lp - lm) / (2 * h)
        for j in range(3):
            print(k, j, grads[j] @ dirn, fd[j])
            assert abs(grads[j] @ dirn - fd[j]) <= 1e-6 * abs(lam[0, j]), (k, j)
        ksp = sg.ks_aggregate(-lp / sig, 100.0)[0]
        ksm = sg.ks_aggregate(-lm / sig, 100.0)[0]
        fdks = (ksp - ksm) / (2 * h)
        print("ks", gks @ dirn, fdks)
        assert abs(gks @ dirn - fdks) <= 1e-6 * max(abs(fdks), 1.0)
    B.destroy()


def test_buckling_coordinate_gradient_matches_oracle_fd():
    cyl = Cylinder(8, 16)
    thick = seeded_thickness(cyl.nelems, 61)
    ctx = _cyl_ctx(cyl, thick)
    nev = 3
    B = ctx.buckling(sigma=0.0, nev=nev, max_lanczos=70, tol=1e-12)
    f = cyl.edge_load(1.0)
    B.solve(force=f)
    lam = B.eigenvalues()[0][0]
    sig = B.info()["sigma"][0]
    coef = np.array([1.0, -0.5, 0.25])
    _, gX = B.eigen_sens(coef, thickness=False, coords=True)
    rng = np.random.default_rng(3001)

    def central(dirn, h):
        Xp = (cyl.Xpts.ravel() + h * dirn).reshape(-1, 3)
        Xm = (cyl.Xpts.ravel() - h * dirn).reshape(-1, 3)
        lp = O.linear_buckling(Xp, cyl.conn, thick[0], cyl.bcs(), f, sig, nev=nev, max_lanczos=70)[0]
        lm = O.linear_buckling(Xm, cyl.conn, thick[0], cyl.bcs(), f, sig, nev=nev, max_lanczos=70)[0]
        return coef @ (lp - lm) / (2 * h)

    for k in range(2):
        # a random nodal perturbation makes the loads strongly curved along
        # it: Richardson-extrapolated central differences (error O(h^4))
        dirn = rng.standard_normal(3 * cyl.nnodes)
        h = 1e-7
        fd = (4.0 * central(dirn, h) - central(dirn, 2 * h)) / 3.0
        print(k, gX[0] @ dirn, fd)
        assert abs(gX[0] @ dirn - fd) <= 1e-6 * abs(fd), k
    B.destroy()


def test_batched_gradient_matches_single_design():
    cyl = Cylinder(8, 16)
    thick = seeded_thickness(cyl.nelems, 41, ndesigns=3)
    f = cyl.edge_load(1.0)
    coef = np.array([1.0, 0.5, 0.0])
    ctx = _cyl_ctx(cyl, thick)
    B = ctx.buckling(sigma=0.0, nev=3, max_lanczos=60, tol=1e-12)
    B.solve(force=f)
    lam_b = B.eigenvalues()[0]
    g_b = B.eigen_sens(coef)[0]
    B.destroy()
    ctx.close()
    for d in range(3):
        c1 = _cyl_ctx(cyl, thick[d:d + 1])
        B1 = c1.buckling(sigma=0.0, nev=3, max_lanczos=60, tol=1e-12)
        B1.solve(force=f)
        lam1 = B1.eigenvalues()[0]
        g1 = B1.eigen_sens(coef)[0]
        assert np.abs(lam1[0] - lam_b[d]).max() <= 1e-10 * lam1[0].max()
        assert np.abs(g1[0] - g_b[d]).max() <= 1e-8 * np.abs(g1[0]).max(), d
        B1.destroy()
        c1.close()
