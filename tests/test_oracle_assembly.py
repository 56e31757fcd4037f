"""Oracle assembly path against brute force: pattern, scatter, BCs, SpMV,
first-appearance numbering."""
import numpy as np

import pyoracle as O
from helpers import random_mesh
from shellgpu.problems import Cylinder, first_appearance


def brute_pattern(nnodes, conn):
    adj = [set() for _ in range(nnodes)]
    for e in conn:
        for a in e:
            adj[a].update(int(b) for b in e)
    rowp = np.zeros(nnodes + 1, np.int32)
    cols = []
    for i in range(nnodes):
        c = sorted(adj[i])
        cols += c
        rowp[i + 1] = len(cols)
    return rowp, np.array(cols, np.int32)


def test_pattern_is_sorted_unique_node_adjacency():
    for conn, X in [random_mesh(5, 4, 1), (Cylinder(6, 8).conn, Cylinder(6, 8).Xpts)]:
        nn = X.shape[0]
        rowp, cols = O.node_csr(nn, conn)
        r2, c2 = brute_pattern(nn, conn)
        assert np.array_equal(rowp, r2) and np.array_equal(cols, c2)


def test_cylinder_pattern_size_matches_closed_form():
    # interior rows: 3x3 stencil, end rows: 2x3 (periodic in theta)
    for na, nc in [(4, 8), (20, 40)]:
        cyl = Cylinder(na, nc)
        rowp, cols = O.node_csr(cyl.nnodes, cyl.conn)
        assert cols.size == 9 * (na - 1) * nc + 6 * 2 * nc


def test_first_appearance_numbering():
    cyl = Cylinder(5, 7)
    # the generator already applies the rule: renumbering again is the identity
    assert np.array_equal(O.first_appearance(cyl.nnodes, cyl.conn), np.arange(cyl.nnodes))
    rng = np.random.default_rng(3)
    perm = rng.permutation(cyl.nnodes)
    c2 = perm[cyl.conn]
    assert np.array_equal(O.first_appearance(cyl.nnodes, c2), first_appearance(c2, cyl.nnodes))


def test_scatter_bcs_matvec_against_dense():
    conn, X = random_mesh(4, 3, 5)
    nn = X.shape[0]
    m = O.material()
    thick = np.linspace(0.002, 0.004, conn.shape[0])
    rng = np.random.default_rng(11)
    u = 1e-3 * rng.standard_normal(6 * nn)
    rowp, cols = O.node_csr(nn, conn)
    vals = O.assemble_mat(O.KIND_KNL, 1, m, X, conn, thick, rowp, cols, u=u)
    dense = np.zeros((6 * nn, 6 * nn))
    for e in range(conn.shape[0]):
        nodes = conn[e]
        q = u.reshape(-1, 6)[nodes].ravel()
        _, Ke = O.elem_jacobian(1, m, thick[e], 0.0, X[nodes].ravel(), q)
        idx = (6 * nodes[:, None] + np.arange(6)).ravel()
        dense[np.ix_(idx, idx)] += Ke
    A = O.bsr_to_dense(rowp, cols, vals)
    assert np.abs(A - dense).max() <= 1e-14 * np.abs(dense).max()
    # residual assembly = sum of element residuals
    R = O.assemble_res(1, m, X, conn, thick, u=u)
    Rd = np.zeros(6 * nn)
    for e in range(conn.shape[0]):
        nodes = conn[e]
        Rd[(6 * nodes[:, None] + np.arange(6)).ravel()] += O.elem_residual(
            1, m, thick[e], 0.0, X[nodes].ravel(), u.reshape(-1, 6)[nodes].ravel())
    assert np.abs(R - Rd).max() <= 1e-14 * np.abs(Rd).max()
    # BCs: masked rows zero, unit diagonal
    bcn = np.array([0, 3], np.int32)
    bcm = np.array([0x3F, 0x05], np.int32)
    vb = O.bsr_apply_bcs(vals, rowp, cols, bcn, bcm)
    Ad = dense.copy()
    for n, msk in zip(bcn, bcm):
        for r in range(6):
            if msk & (1 << r):
                Ad[6 * n + r, :] = 0.0
                Ad[6 * n + r, 6 * n + r] = 1.0
    assert np.array_equal(O.bsr_to_dense(rowp, cols, vb) != 0, Ad != 0)
    assert np.abs(O.bsr_to_dense(rowp, cols, vb) - Ad).max() <= 1e-14 * np.abs(Ad).max()
    x = rng.standard_normal(6 * nn)
    assert np.abs(O.bsr_matvec(rowp, cols, vb, x) - Ad @ x).max() <= 1e-13 * np.abs(Ad @ x).max()


def test_rigid_translation_in_nullspace_of_cylinder_K():
    cyl = Cylinder(4, 12)
    m = O.material()
    thick = np.full(cyl.nelems, 0.002)
    rowp, cols = O.node_csr(cyl.nnodes, cyl.conn)
    vals = O.assemble_mat(O.KIND_K, 0, m, cyl.Xpts, cyl.conn, thick, rowp, cols)
    for d in range(3):
        r = np.zeros(6 * cyl.nnodes)
        r[d::6] = 1.0
        y = O.bsr_matvec(rowp, cols, vals, r)
        assert np.abs(y).max() <= 1e-8 * np.abs(vals).max()


def test_first_appearance_known_answer():
    # TACSCreator.cpp:1150-1202 on one partition: nodes numbered in order
    # of first appearance, elements in order, nodes within an element in
    # order; the result maps old -> new (unreferenced nodes: -1)
    conn = np.array([[5, 2, 7, 1], [2, 9, 1, 3], [0, 5, 3, 9]], np.int32)
    out = O.first_appearance(10, conn)
    expect = np.full(10, -1)
    for old, new in {5: 0, 2: 1, 7: 2, 1: 3, 9: 4, 3: 5, 0: 6}.items():
        expect[old] = new
    assert np.array_equal(out, expect)
