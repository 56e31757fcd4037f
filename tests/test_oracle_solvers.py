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
