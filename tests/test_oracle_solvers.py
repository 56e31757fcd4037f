"""Oracle solve path pinned against dense linear algebra and closed forms."""
import numpy as np
import scipy.linalg as sl

import pyoracle as O
from shellgpu.problems import Cylinder


def test_tridiagonal_ql_matches_toeplitz_closed_form():
    for n in (1, 2, 7, 40):
        d = 2.0 * np.ones(n)
        e = -np.ones(n - 1)
        w, z = O.tridiag_eig(d, e)
        exact = 2.0 - 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
        assert np.abs(w - exact).max() < 1e-13
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        for k in range(n):
            assert np.linalg.norm(T @ z[k] - w[k] * z[k]) < 1e-12


def test_buckling_matches_dense_generalized_eigensolve():
    cyl = Cylinder(4, 8)
    thick = np.random.default_rng(5).uniform(0.0015, 0.0025, cyl.nelems)
    sigma = 0.9 * cyl.classical_load()
    lam, phi, path, resid, ns = O.linear_buckling(cyl.Xpts, cyl.conn, thick, cyl.bcs(), cyl.edge_load(1.0), sigma,
                                                  nev=5, max_lanczos=80)
    rowp, cols = O.node_csr(cyl.nnodes, cyl.conn)
    m = O.material()
    K = O.bsr_to_dense(rowp, cols, O.assemble_mat(O.KIND_K, 0, m, cyl.Xpts, cyl.conn, thick, rowp, cols))
    bn, bm = cyl.bcs()
    fixed = {6 * b + r for b, mk in zip(bn, bm) for r in range(6) if mk & (1 << r)}
    free = np.array(sorted(set(range(6 * cyl.nnodes)) - fixed))
    f = cyl.edge_load(1.0)
    u = np.zeros(6 * cyl.nnodes)
    u[free] = np.linalg.solve(K[np.ix_(free, free)], f[free])
    assert np.abs(u - path).max() <= 1e-8 * np.abs(u).max()
    G = O.bsr_to_dense(rowp, cols, O.assemble_mat(O.KIND_G, 1, m, cyl.Xpts, cyl.conn, thick, rowp, cols, u=u))
    w = sl.eig(K[np.ix_(free, free)], -G[np.ix_(free, free)], right=False)
    w = np.real(w[np.isfinite(w)])
    w = w[np.argsort(np.abs(w))][:5]
    assert np.abs(lam - w).max() <= 1e-9 * np.abs(w).max()
    assert resid.max() < 1e-8


def _dense_buckling(cyl, thick, u):
    rowp, cols = O.node_csr(cyl.nnodes, cyl.conn)
    m = O.material()
    K = O.bsr_to_dense(rowp, cols, O.assemble_mat(O.KIND_K, 0, m, cyl.Xpts, cyl.conn, thick, rowp, cols))
    G = O.bsr_to_dense(rowp, cols, O.assemble_mat(O.KIND_G, 1, m, cyl.Xpts, cyl.conn, thick, rowp, cols, u=u))
    return K, G


def test_buckling_with_prescribed_displacements_and_given_path():
    cyl = Cylinder(4, 8)
    thick = np.random.default_rng(6).uniform(0.0015, 0.0025, cyl.nelems)
    ring0, ring1 = cyl.ring_nodes(0), cyl.ring_nodes(cyl.na)
    bn = np.concatenate([ring0, ring1]).astype(np.int32)
    bm = np.full(bn.size, 0x3F, np.int32)
    vals = np.zeros((bn.size, 6))
    vals[ring0.size:, 0] = -1e-5  # axial end shortening, displacement control
    fixed = np.array(sorted({6 * b + r for b in bn for r in range(6)}))
    free = np.array(sorted(set(range(6 * cyl.nnodes)) - set(fixed)))
    ubar = np.zeros(6 * cyl.nnodes)
    for b, v in zip(bn, vals):
        ubar[6 * b:6 * b + 6] = v
    K, _ = _dense_buckling(cyl, thick, None)
    u = ubar.copy()
    u[free] = np.linalg.solve(K[np.ix_(free, free)], -K[np.ix_(free, fixed)] @ ubar[fixed])
    sigma = 100.0
    lam, _, path, _, _ = O.linear_buckling(cyl.Xpts, cyl.conn, thick, (bn, bm, vals.ravel()),
                                           np.zeros(6 * cyl.nnodes), sigma, nev=4, max_lanczos=80)
    assert np.abs(path - u).max() <= 1e-9 * np.abs(u).max()
    _, G = _dense_buckling(cyl, thick, u)
    w = sl.eig(K[np.ix_(free, free)], -G[np.ix_(free, free)], right=False)
    w = np.real(w[np.isfinite(w)])
    w = w[np.argsort(np.abs(w))][:4]
    assert np.abs(lam - w).max() <= 1e-9 * np.abs(w).max()

    # a given path (TACSLinearBuckling::solve with u0): BC values imposed on it
    u0 = 3.0 * u + 1e-7 * np.random.default_rng(7).standard_normal(u.size)
    lam0, _, path0, _, _ = O.linear_buckling(cyl.Xpts, cyl.conn, thick, (bn, bm, vals.ravel()),
                                             np.zeros(6 * cyl.nnodes), sigma, nev=4, max_lanczos=80, u0=u0)
    ref0 = u0.copy()
    ref0[fixed] = ubar[fixed]
    assert np.array_equal(path0, ref0)
    _, G0 = _dense_buckling(cyl, thick, ref0)
    w0 = sl.eig(K[np.ix_(free, free)], -G0[np.ix_(free, free)], right=False)
    w0 = np.real(w0[np.isfinite(w0)])
    w0 = w0[np.argsort(np.abs(w0))][:4]
    assert np.abs(lam0 - w0).max() <= 1e-9 * np.abs(w0).max()


def _bandwidth(rowp, cols, perm):
    """Largest |pos(i) - pos(j)| over the pattern for the order perm
    (perm[k] = node at position k)."""
    pos = np.empty(perm.size, np.int64)
    pos[perm] = np.arange(perm.size)
    bw = 0
    for i in range(rowp.size - 1):
        for j in cols[rowp[i]:rowp[i + 1]]:
            bw = max(bw, abs(int(pos[i]) - int(pos[j])))
    return bw


def test_rcm_is_a_permutation_of_small_bandwidth():
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    # a path graph with shuffled labels: any Cuthill-McKee order from an
    # end walks the path, bandwidth 1
    n = 23
    rng = np.random.default_rng(7)
    lab = rng.permutation(n)
    edges = [(lab[k], lab[k + 1]) for k in range(n - 1)]
    adj = [{i} for i in range(n)]
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    rowp = np.zeros(n + 1, np.int32)
    cols = []
    for i in range(n):
        cols += sorted(adj[i])
        rowp[i + 1] = len(cols)
    cols = np.array(cols, np.int32)
    perm = O.rcm(rowp, cols)
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert _bandwidth(rowp, cols, perm) == 1
    # the cylinder: level sets are rings of nc nodes (bandwidth about one
    # ring plus the periodic seam), and no worse than scipy's RCM
    for na, nc in [(6, 10), (12, 8)]:
        cyl = Cylinder(na, nc)
        rowp, cols = O.node_csr(cyl.nnodes, cyl.conn)
        perm = O.rcm(rowp, cols)
        assert np.array_equal(np.sort(perm), np.arange(cyl.nnodes))
        bw = _bandwidth(rowp, cols, perm)
        A = sp.csr_matrix((np.ones(cols.size), cols, rowp), shape=(cyl.nnodes, cyl.nnodes))
        ref = _bandwidth(rowp, cols, np.asarray(reverse_cuthill_mckee(A, symmetric_mode=True)))
        assert bw <= ref and bw <= 2 * nc, (bw, ref)


def test_band_solve_matches_dense_solve():
    # a nonsymmetric, block-diagonally dominant matrix on a mesh pattern:
    # the band LU without pivoting is stable on it
    cyl = Cylinder(5, 6)
    rowp, cols = O.node_csr(cyl.nnodes, cyl.conn)
    rng = np.random.default_rng(12)
    vals = rng.uniform(-1.0, 1.0, (cols.size, 6, 6))
    for i in range(cyl.nnodes):
        for p in range(rowp[i], rowp[i + 1]):
            if cols[p] == i:
                vals[p] += 60.0 * np.eye(6)
    A = O.bsr_to_dense(rowp, cols, vals.reshape(-1))
    b = rng.standard_normal(6 * cyl.nnodes)
    x = O.band_solve(rowp, cols, vals.reshape(-1), b)
    xd = np.linalg.solve(A, b)
    assert np.abs(x - xd).max() <= 1e-12 * np.abs(xd).max()
