"""Assembler and matrix entry points beyond the buckling path's own, through
the C ABI, against the oracle:

* assembleJacobian(alpha, beta, gamma) (TACSAssembler.cpp:4084-4178 with
  TACSShellElement::addJacobian, TACSShellElement.h:455-672): R = R(u) +
  M u'' and A = alpha dR/du + gamma M, BCs applied, for both element models,
  from one element pass; beta and du/dt do not enter (the linearized-rotation
  director has no velocity terms); the matrix overwrites what A held;
* assembleMatCombo (TACSAssembler.cpp:4264-4310): a weighted sum of K, G(u)
  and M with BCs applied once;
* addJacobianVecProduct (TACSAssembler.cpp:4334-4390): y += s J x and the
  transposed product, J = alpha dR/du + gamma M from element matrices, then
  y[bc] = 0;
* TACSMat::multTranspose, scale, addDiag and addValues (KSM.h:153-180) against
  dense products on random block values, including repeated, negative and
  out-of-pattern blocks;
* psi^T (dA/dt) phi, d(psi^T A phi)/dX and d(psi^T G phi)/du with psi != phi
  against five-point differences of the oracle's element matrices (and G's
  affinity in the path).
"""
import numpy as np
import pytest

import pyoracle as O
import shellgpu as sg
from helpers import bsr_row_floor_ratio, random_mesh

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _ctx(conn, X, thick, bcs=None):
    ctx = sg.Context(0)
    ctx.set_mesh(conn, X)
    ctx.set_material(70e9, 0.3, 2700.0, 0.0, 10.0)
    if bcs is not None:
        ctx.set_bcs(*bcs)
    ctx.set_designs(thick)
    return ctx


def _bcs():
    return np.array([0, 1, 7], np.int32), np.array([0x3F, 0x3F, 0x15], np.int32)


def _bc_mask(nn, bcs):
    m = np.zeros((nn, 6), bool)
    for node, mask in zip(*bcs):
        for k in range(6):
            if mask >> k & 1:
                m[node, k] = True
    return m.ravel()


@pytest.mark.parametrize("model", [sg.LINEAR, sg.NONLINEAR])
def test_assemble_jacobian_alpha_beta_gamma(model):
    conn, X = random_mesh(5, 4, seed=31)
    nn = X.shape[0]
    bcs = _bcs()
    rng = np.random.default_rng(32)
    nb = 2
    thick = rng.uniform(0.0015, 0.0025, (nb, conn.shape[0]))
    u = 1e-3 * rng.standard_normal((nb, 6 * nn))
    udd = 50.0 * rng.standard_normal((nb, 6 * nn))
    alpha, beta, gamma = 0.7, 0.3, 4.0e7  # gamma M of the size of alpha K
    m = O.material()
    with _ctx(conn, X, thick, bcs) as ctx:
        rowp, cols = ctx.pattern()
        du, dudd, dudot = ctx.vector(u), ctx.vector(udd), ctx.vector(rng.standard_normal((nb, 6 * nn)))
        R, A = ctx.vector(), ctx.matrix()
        A.set_values(1e3 * rng.standard_normal((nb, ctx.nnzb, 6, 6)))  # overwritten
        ctx.assemble_jacobian(A, R, du, model=model, alpha=alpha, beta=beta, gamma=gamma, dudt=dudot, d2udt2=dudd)
        Ag, Rg = A.values(), R.download().reshape(nb, -1)
        # residual alone, and the matrix without a residual (the other kernels)
        R2, A2 = ctx.vector(), ctx.matrix()
        ctx.assemble_res(R2, du, model=model)
        ctx.assemble_jacobian(A2, None, du, model=model, alpha=alpha)
        R2, A2 = R2.download().reshape(nb, -1), A2.values()
    bc = _bc_mask(nn, bcs)
    for d in range(nb):
        kind = O.KIND_KNL if model == sg.NONLINEAR else O.KIND_K
        K = O.assemble_mat(kind, model, m, X, conn, thick[d], rowp, cols, u=u[d])
        M = O.assemble_mat(O.KIND_M, 0, m, X, conn, thick[d], rowp, cols)
        Aref = O.bsr_apply_bcs(alpha * K + gamma * M, rowp, cols, *bcs)
        assert bsr_row_floor_ratio(rowp, Ag[d], Aref) <= TOL, d
        Rref = O.assemble_res(model, m, X, conn, thick[d], u=u[d]) + O.bsr_matvec(rowp, cols, M, udd[d])
        Rref[bc] = u[d][bc]  # R[bc] = u - ubar, ubar = 0
        assert np.abs(Rg[d] - Rref).max() <= TOL * np.abs(Rref).max(), d
        # the fused pass against the separate residual and matrix kernels
        R2ref = O.assemble_res(model, m, X, conn, thick[d], u=u[d])
        R2ref[bc] = u[d][bc]
        assert np.abs(R2[d] - R2ref).max() <= TOL * np.abs(R2ref).max(), d
        assert bsr_row_floor_ratio(rowp, A2[d], O.bsr_apply_bcs(alpha * K, rowp, cols, *bcs)) <= TOL, d


def test_assemble_mat_combo():
    conn, X = random_mesh(4, 5, seed=41)
    nn = X.shape[0]
    bcs = _bcs()
    rng = np.random.default_rng(42)
    thick = rng.uniform(0.0015, 0.0025, (1, conn.shape[0]))
    u = 1e-3 * rng.standard_normal(6 * nn)
    m = O.material()
    types, scale = [sg.STIFFNESS, sg.GEOMETRIC, sg.MASS], [1.0, 0.5, -3.0e7]
    with _ctx(conn, X, thick, bcs) as ctx:
        rowp, cols = ctx.pattern()
        A = ctx.matrix()
        A.set_values(rng.standard_normal((1, ctx.nnzb, 6, 6)))  # zeroed first
        ctx.assemble_mat_combo(A, types, scale, ctx.vector(u))
        Ag = A.values()[0]
    K = O.assemble_mat(O.KIND_K, 0, m, X, conn, thick[0], rowp, cols)
    G = O.assemble_mat(O.KIND_G, 1, m, X, conn, thick[0], rowp, cols, u=u)
    M = O.assemble_mat(O.KIND_M, 0, m, X, conn, thick[0], rowp, cols)
    ref = O.bsr_apply_bcs(K + 0.5 * G - 3.0e7 * M, rowp, cols, *bcs)
    assert bsr_row_floor_ratio(rowp, Ag, ref) <= TOL


@pytest.mark.parametrize("model", [sg.LINEAR, sg.NONLINEAR])
@pytest.mark.parametrize("transpose", [False, True])
def test_jacobian_vec_product(model, transpose):
    conn, X = random_mesh(5, 3, seed=51)
    nn = X.shape[0]
    bcs = _bcs()
    rng = np.random.default_rng(52)
    nb = 2
    thick = rng.uniform(0.0015, 0.0025, (nb, conn.shape[0]))
    u = 1e-3 * rng.standard_normal((nb, 6 * nn))
    x = rng.standard_normal((nb, 6 * nn))
    y0 = rng.standard_normal((nb, 6 * nn))
    s, alpha, gamma = -1.5, 0.8, 2.0e7
    m = O.material()
    with _ctx(conn, X, thick, bcs) as ctx:
        rowp, cols = ctx.pattern()
        dy = ctx.vector(y0)
        ctx.add_jacobian_vec_product(dy, ctx.vector(x), scale=s, alpha=alpha, gamma=gamma, u=ctx.vector(u),
                                     model=model, transpose=transpose)
        y = dy.download().reshape(nb, -1)
    bc = _bc_mask(nn, bcs)
    for d in range(nb):
        kind = O.KIND_KNL if model == sg.NONLINEAR else O.KIND_K
        J = alpha * O.assemble_mat(kind, model, m, X, conn, thick[d], rowp, cols, u=u[d]) + \
            gamma * O.assemble_mat(O.KIND_M, 0, m, X, conn, thick[d], rowp, cols)
        Jd = O.bsr_to_dense(rowp, cols, J)  # no BCs on J
        ref = y0[d] + s * ((Jd.T if transpose else Jd) @ x[d])
        ref[bc] = 0.0
        assert np.abs(y[d] - ref).max() <= TOL * np.abs(ref).max(), d


def test_matrix_operations():
    conn, X = random_mesh(4, 3, seed=61)
    nn = X.shape[0]
    rng = np.random.default_rng(62)
    nb = 2
    thick = rng.uniform(0.0015, 0.0025, (nb, conn.shape[0]))
    with _ctx(conn, X, thick) as ctx:
        rowp, cols = ctx.pattern()
        vals = rng.standard_normal((nb, ctx.nnzb, 6, 6))  # not symmetric
        A = ctx.matrix()
        A.set_values(vals)
        x = rng.standard_normal((nb, 6 * nn))
        dx, dy = ctx.vector(x), ctx.vector()
        A.mult_transpose(dx, dy)
        yt = dy.download().reshape(nb, -1)
        for d in range(nb):
            D = O.bsr_to_dense(rowp, cols, vals[d])
            assert np.abs(yt[d] - D.T @ x[d]).max() <= 1e-13 * np.abs(D).sum(axis=0).max() * np.abs(x).max()
        A.scale(-2.5)
        A.add_diag(7.0)
        ref = -2.5 * vals
        diag = np.array([rowp[i] + np.searchsorted(cols[rowp[i]:rowp[i + 1]], i) for i in range(nn)])
        ref[:, diag] += 7.0 * np.eye(6)
        assert np.array_equal(A.values(), ref)
        # addValues: repeated blocks summed in input order, negative indices
        # skipped, an out-of-pattern block reported after the others are added
        nblk = 40
        br = rng.integers(0, nn, nblk)
        bcl = np.array([cols[rng.integers(rowp[r], rowp[r + 1])] for r in br])
        br[5], bcl[5] = br[3], bcl[3]  # a repeat
        br[7] = -1  # skipped
        bv = rng.standard_normal((nblk, 6, 6))
        A.add_blocks(br, bcl, bv, design=1)
        exp = ref.copy()
        for k in range(nblk):
            if br[k] < 0:
                continue
            p = rowp[br[k]] + np.searchsorted(cols[rowp[br[k]]:rowp[br[k] + 1]], bcl[k])
            exp[1, p] += bv[k]
        assert np.abs(A.values() - exp).max() <= 1e-14 * np.abs(exp).max()
        far = [r for r in range(nn) if 0 not in cols[rowp[r]:rowp[r + 1]]][0]
        with pytest.raises(sg.ShellGPUError, match="not in the pattern"):
            A.add_blocks([far, 0], [0, 0], np.ones((2, 6, 6)), design=0)
        exp[0, rowp[0]] += 1.0  # block (0, 0) is the first of row 0
        assert np.abs(A.values() - exp).max() <= 1e-14 * np.abs(exp).max()


def _five_point(f, t, h):
    return (f(t - 2 * h) - 8 * f(t - h) + 8 * f(t + h) - f(t + 2 * h)) / (12 * h)


def test_inner_products_with_psi_not_phi():
    conn, X = random_mesh(4, 3, seed=71)
    nn, ne = X.shape[0], conn.shape[0]
    rng = np.random.default_rng(72)
    thick = rng.uniform(0.002, 0.02, (1, ne))
    psi, phi = rng.standard_normal(6 * nn), rng.standard_normal(6 * nn)
    u = 1e-3 * rng.standard_normal(6 * nn)
    m = O.material()
    with _ctx(conn, X, thick) as ctx:
        dpsi, dphi, du = ctx.vector(psi), ctx.vector(phi), ctx.vector(u)
        out = {}
        for name, kind in (("K", sg.STIFFNESS), ("G", sg.GEOMETRIC), ("M", sg.MASS)):
            g = ctx.design_vector()
            ctx.add_mat_dv_sens_inner_product(g, kind, dphi, du, scale=1.0, psi=dpsi)
            out[name] = g.download()
        gX = {}
        for name, kind in (("K", sg.STIFFNESS), ("G", sg.GEOMETRIC)):
            g = ctx.coord_vector()
            ctx.add_mat_xpt_sens_inner_product(g, kind, dphi, du, scale=1.0, psi=dpsi)
            gX[name] = g.download()
        gu = ctx.vector()
        ctx.eval_mat_sv_sens_inner_product(dphi, gu, psi=dpsi)
        gu = gu.download()

    def emat(name, t, Xe, ue):
        if name == "K":
            return O.elem_stiffness(0, m, t, 0.0, Xe, np.zeros(24))
        if name == "G":
            return O.elem_geometric_step(m, t, 0.0, Xe, ue, 1.0)
        return O.elem_mass(m, t, Xe)

    for name in ("K", "G", "M"):
        ref = np.zeros(ne)
        for e in range(ne):
            Xe = X[conn[e]].ravel()
            idx = (6 * conn[e][:, None] + np.arange(6)).ravel()
            t = thick[0, e]
            ref[e] = _five_point(lambda s: psi[idx] @ emat(name, s, Xe, u[idx]) @ phi[idx], t, 0.05 * t)
        err = np.abs(out[name] - ref).max() / np.abs(ref).max()
        print("dt", name, err)
        assert err <= 1e-10, name

    def total(name, Xp):
        s = 0.0
        for e in range(ne):
            Xe = Xp[conn[e]].ravel()
            idx = (6 * conn[e][:, None] + np.arange(6)).ravel()
            s += psi[idx] @ emat(name, thick[0, e], Xe, u[idx]) @ phi[idx]
        return s

    h = 3e-4
    for name in ("K", "G"):
        ref = np.zeros(3 * nn)
        for k in range(3 * nn):
            def f(x, k=k):
                Xp = X.copy().ravel()
                Xp[k] = x
                return total(name, Xp.reshape(-1, 3))
            ref[k] = _five_point(f, X.ravel()[k], h)
        err = np.abs(gX[name] - ref).max() / np.abs(ref).max()
        print("dX", name, err)
        assert err <= 1e-8, name
    # d(psi^T G(u) phi)/du: G is affine in the path
    rowp, cols = O.node_csr(nn, conn)

    def pGp(w):
        G = O.assemble_mat(O.KIND_G, 0, m, X, conn, thick[0], rowp, cols, u=w)
        return psi @ O.bsr_matvec(rowp, cols, G, phi)

    base = pGp(np.zeros(6 * nn))
    for k in range(3):
        w = rng.standard_normal(6 * nn)
        ref = pGp(w) - base
        assert abs(gu @ w - ref) <= 1e-9 * max(abs(ref), 1e-30), k


@pytest.mark.parametrize("nb", [11, 16])
def test_mult_identical_designs(nb):
    """Load cases of one structure: products with K, M and the linear-model
    Jacobian read one copy for every design's vector; the result equals the
    per-design product bit for bit, over groups of eight vectors with and
    without a ragged remainder, and a per-design change (addValues) or a state-
    dependent type (G(u), the nonlinear tangent) turns the shared read off."""
    conn, X = random_mesh(5, 4, seed=71)
    nn = X.shape[0]
    rng = np.random.default_rng(72)
    thick = np.tile(rng.uniform(0.0015, 0.0025, conn.shape[0]), (nb, 1))
    with _ctx(conn, X, thick, _bcs()) as ctx:
        A, P = ctx.matrix(), ctx.matrix()
        x = rng.standard_normal((nb, 6 * nn))
        u = 1e-3 * rng.standard_normal((nb, 6 * nn))
        dx, dy, dz, du = ctx.vector(x), ctx.vector(), ctx.vector(), ctx.vector(u)
        R = ctx.vector()

        def check(exact=True):
            A.mult(dx, dy)
            P.set_values(A.values())  # the same values, read per design
            P.mult(dx, dz)
            y, z = dy.download(), dz.download()
            if exact:
                assert np.array_equal(y, z)
            return y.reshape(nb, -1), A.values()

        for mattype in (sg.STIFFNESS, sg.MASS):
            ctx.assemble_mattype(mattype, A, du)
            check()
        ctx.assemble_mattype(sg.GEOMETRIC, A, du)
        check()
        ctx.assemble_mattype(sg.TANGENT_NONLINEAR, A, du)
        check()
        ctx.assemble_jacobian(A, R, du, model=sg.LINEAR, alpha=1.5, gamma=-0.25)
        check()
        ctx.assemble_mat_combo(A, [sg.STIFFNESS, sg.MASS], [1.0, 2.0], du)
        check()
        A.scale(-2.0)
        A.add_diag(3.0)
        check()
        ctx.assemble_mattype(sg.STIFFNESS, A)
        y0, _ = check()
        rowp, cols = ctx.pattern()
        A.add_blocks([0], [0], np.ones((1, 6, 6)), design=4)
        y1, v1 = check()
        for d in (0, 4):
            D = O.bsr_to_dense(rowp, cols, v1[d])
            assert np.abs(y1[d] - D @ x[d]).max() <= 1e-13 * np.abs(D).sum(axis=1).max() * np.abs(x).max()
        assert np.array_equal(y1[0], y0[0]) and not np.array_equal(y1[4], y0[4])
