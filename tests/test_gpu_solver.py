"""GPU direct solver, GMRES and linear buckling against the oracle: static
path, eigenvalues within 1e-8 relative, eigenvectors by principal angles of
each (near-)degenerate cluster within 1e-6 in the K inner product, and
batched designs against per-design oracle runs."""
import numpy as np
import pytest

import pyoracle as O
import shellgpu as sg
from shellgpu.problems import Cylinder

pytestmark = pytest.mark.gpu


def set_factor(monkeypatch, factor):
    """cr: block cyclic reduction; thomas: the sequential level factorisation;
    -block: with block Lanczos (blocks of 4 vectors; cyclic reduction then
    solves the 4 vectors of a block together)"""
    monkeypatch.setenv("SHELLGPU_FACTOR", factor.split("-")[0])
    if factor.endswith("-block"):
        monkeypatch.setenv("SHELLGPU_LANCZOS_BLOCK", "4")


def make_ctx(cyl, thick):
    ctx = sg.Context(0)
    ctx.set_mesh(cyl.conn, cyl.Xpts)
    ctx.set_material(70e9, 0.3)
    ctx.set_bcs(*cyl.bcs())
    ctx.set_designs(thick)
    ctx.set_ordering_seed(cyl.ring_nodes(0))
    return ctx


@pytest.mark.parametrize("factor", ["cr", "thomas"])
def test_direct_solve_and_gmres(factor, monkeypatch):
    # cr: block cyclic reduction over the (uniform) ring levels;
    # thomas: the sequential level Cholesky
    monkeypatch.setenv("SHELLGPU_FACTOR", factor)
    cyl = Cylinder(10, 20)
    thick = np.random.default_rng(1).uniform(0.0015, 0.0025, cyl.nelems)
    ctx = make_ctx(cyl, thick)
    K = ctx.matrix()
    ctx.assemble_mattype(sg.STIFFNESS, K)
    F = ctx.factor()
    nl, mx, _ = F.info()
    assert nl == cyl.na + 1 and mx == 6 * cyl.nc  # one level per ring
    F.compute(K)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(6 * cyl.nnodes)
    db, dx = ctx.vector(b), ctx.vector()
    its, rn = ctx.gmres(K, db, dx, F=F, rtol=1e-12)
    x = dx.download()
    rowp, cols = ctx.pattern()
    Kref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_K, 0, O.material(), cyl.Xpts, cyl.conn, thick, rowp, cols),
                           rowp, cols, *cyl.bcs())
    r = O.bsr_matvec(rowp, cols, Kref, x) - b
    # explicit level inverses: the true residual floor is ~cond * eps
    assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(b)
    xr = O.band_solve(rowp, cols, Kref, b)
    assert np.abs(x - xr).max() <= 1e-8 * np.abs(xr).max()


def _subspace_angle(V, W, rowp, cols, Kvals):
    KW = np.stack([O.bsr_matvec(rowp, cols, Kvals, w) for w in W])
    M = V @ KW.T
    s = np.linalg.svd(M, compute_uv=False)
    return float(np.arccos(min(1.0, s.min())))


@pytest.mark.parametrize("factor", ["cr", "thomas", "thomas-block", "cr-block"])
@pytest.mark.parametrize("na,nc", [(20, 40)])
def test_buckling_matches_oracle(na, nc, factor, monkeypatch):
    set_factor(monkeypatch, factor)
    cyl = Cylinder(na, nc)
    thick = np.full(cyl.nelems, 0.002)
    sigma = 0.9 * cyl.classical_load()
    ctx = make_ctx(cyl, thick)
    nev = 8  # 5 wanted + margin so that a degenerate pair at the boundary is whole
    out = ctx.linear_buckling(sigma, cyl.edge_load(1.0), nev=nev, max_lanczos=120, tol=1e-12, want_vectors=True,
                              want_path=True)
    lam, phi, path, resid, ns = O.linear_buckling(cyl.Xpts, cyl.conn, thick, cyl.bcs(), cyl.edge_load(1.0), sigma,
                                                  nev=nev, max_lanczos=120)
    g = out["lambda"][0]
    print("gpu", g, "oracle", lam, "steps", out["steps"])
    assert np.abs(g[:5] - lam[:5]).max() <= 1e-8 * np.abs(lam[:5]).max()
    assert np.abs(out["path"][0] - path).max() <= 1e-8 * np.abs(path).max()
    rowp, cols = ctx.pattern()
    Kref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_K, 0, O.material(), cyl.Xpts, cyl.conn, thick, rowp, cols),
                           rowp, cols, *cyl.bcs())
    # clusters of (near-)degenerate eigenvalues compared as subspaces; the
    # clusters covering the first 5 eigenvalues must close before index nev
    i = 0
    while i < 5:
        j = i + 1
        while j < nev and abs(lam[j] - lam[i]) <= 1e-6 * abs(lam[i]):
            j += 1
        assert j < nev, "cluster reaches the end of the computed set"
        ang = _subspace_angle(out["phi"][0][i:j], phi[i:j], rowp, cols, Kref)
        print("cluster", i, j, "angle", ang)
        assert ang <= 1e-6
        i = j


@pytest.mark.parametrize("factor", ["cr", "thomas", "thomas-block"])
def test_batched_designs_match_single_designs(factor, monkeypatch):
    set_factor(monkeypatch, factor)
    cyl = Cylinder(8, 16, imperfection=(2e-4, 2, 3))
    rng = np.random.default_rng(9)
    thick = rng.uniform(0.0015, 0.0025, (3, cyl.nelems))
    sigma = 0.8 * cyl.classical_load()
    ctx = make_ctx(cyl, thick)
    out = ctx.linear_buckling(sigma, cyl.edge_load(1.0), nev=4, max_lanczos=80, tol=1e-12)
    for b in range(3):
        lam, _, _, _, _ = O.linear_buckling(cyl.Xpts, cyl.conn, thick[b], cyl.bcs(), cyl.edge_load(1.0), sigma,
                                            nev=4, max_lanczos=80)
        assert np.abs(out["lambda"][b] - lam).max() <= 1e-8 * np.abs(lam).max()


def test_batched_designs_converging_at_different_steps(monkeypatch):
    """Designs of one batch whose wanted pairs converge at different Lanczos
    steps (thickness x1, x1.4, x2 under one shift): with the sequential
    factorisation each design stops solving at its own convergence, and its
    loads and Ritz residuals still match the oracle's and its own
    single-design run's."""
    set_factor(monkeypatch, "thomas")
    cyl = Cylinder(8, 16, imperfection=(2e-4, 2, 3))
    base = np.random.default_rng(29).uniform(0.0015, 0.0025, cyl.nelems)
    thick = np.stack([base, 1.4 * base, 2.0 * base])
    sigma = 0.8 * cyl.classical_load()
    f = cyl.edge_load(1.0)
    ctx = make_ctx(cyl, thick)
    out = ctx.linear_buckling(sigma, f, nev=4, max_lanczos=120, tol=1e-12)
    ctx.close()
    steps = []
    for b in range(3):
        c1 = make_ctx(cyl, thick[b])
        one = c1.linear_buckling(sigma, f, nev=4, max_lanczos=120, tol=1e-12)
        c1.close()
        steps.append(one["steps"])
        lam = O.linear_buckling(cyl.Xpts, cyl.conn, thick[b], cyl.bcs(), f, sigma, nev=4, max_lanczos=120)[0]
        print("design", b, "batch", out["lambda"][b], "single", one["lambda"][0], "oracle", lam, "steps", one["steps"])
        assert np.abs(out["lambda"][b] - lam).max() <= 1e-8 * np.abs(lam).max()
        assert np.abs(out["lambda"][b] - one["lambda"][0]).max() <= 1e-10 * np.abs(lam).max()
        assert out["resid"][b].max() <= 1e-12
    # the designs alone converge at different steps, so the batch froze some
    assert len(set(steps)) > 1, steps


@pytest.mark.parametrize("factor", ["cr", "thomas"])
def test_solve_graph_replays_match_direct_launches(factor, monkeypatch):
    """The captured solve graph (SHELLGPU_GRAPH=1), replayed with eager work
    between replays and after a refactorisation with new values, gives
    bit-identical results to the same solve launched kernel by kernel."""
    set_factor(monkeypatch, factor)
    cyl = Cylinder(10, 20)
    rng = np.random.default_rng(11)
    thick = rng.uniform(0.0015, 0.0025, (2, cyl.nelems))
    ctx = make_ctx(cyl, thick)
    K = ctx.matrix()
    ctx.assemble_mattype(sg.STIFFNESS, K)
    n = 2 * 6 * cyl.nnodes
    # a factor decides at its first solve: capture (graph) or direct launches
    monkeypatch.setenv("SHELLGPU_GRAPH", "1")
    Fg = ctx.factor()
    Fg.compute(K)
    Fg.apply(ctx.vector(rng.standard_normal(n)), ctx.vector())  # captures the graph
    monkeypatch.setenv("SHELLGPU_GRAPH", "0")
    Fd = ctx.factor()
    Fd.compute(K)
    for rep in range(6):
        if rep == 3:  # new values in the same buffers: the graph reads them at run time
            ctx.set_designs(thick * 1.1)
            ctx.assemble_mattype(sg.STIFFNESS, K)
            Fg.compute(K)
            Fd.compute(K)
        db = ctx.vector(rng.standard_normal(n))
        xg, xd = ctx.vector(), ctx.vector()
        Fg.apply(db, xg)
        ctx.gmres(K, db, xd, F=Fd, rtol=1e-6)  # eager work between the replays
        Fd.apply(db, xd)
        assert np.array_equal(xg.download(), xd.download()), rep
    Fg.destroy()
    Fd.destroy()
    ctx.close()


def test_gmres_is_bit_reproducible():
    # the batched dot products reduce in a fixed order (two passes, no
    # atomics): two identical solves agree bit for bit
    cyl = Cylinder(10, 20)
    thick = np.random.default_rng(4).uniform(0.0015, 0.0025, (3, cyl.nelems))
    with make_ctx(cyl, thick) as ctx:
        K = ctx.matrix()
        ctx.assemble_mattype(sg.STIFFNESS, K)
        F = ctx.factor()
        F.compute(K)
        b = ctx.vector(np.random.default_rng(5).standard_normal((3, 6 * cyl.nnodes)))
        out = []
        for _ in range(2):
            x = ctx.vector()
            r = ctx.gmres(K, b, x, F=F, msub=10, nrestart=3, rtol=1e-14)
            out.append((r, x.download()))
        assert out[0][0] == out[1][0]
        assert np.array_equal(out[0][1].view(np.int64), out[1][1].view(np.int64))
        # unpreconditioned: many dot products per solve
        out = []
        for _ in range(2):
            x = ctx.vector()
            r = ctx.gmres(K, b, x, msub=40, nrestart=0)
            out.append((r, x.download()))
        assert out[0][0] == out[1][0]
        assert np.array_equal(out[0][1].view(np.int64), out[1][1].view(np.int64))
        K.destroy()
        F.destroy()
