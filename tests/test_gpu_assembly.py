"""GPU assembly parity against the oracle (reference restatement) on the same
inputs: pattern bit-exact, K / nonlinear tangent / residual within 1e-10
(row-floor relative metric), G exact vs the reference's central-difference G,
the FD-replica G, SpMV, and batched designs."""
import numpy as np
import pytest

import pyoracle as O
import shellgpu as sg
from helpers import bsr_row_floor_ratio, random_mesh
from shellgpu.problems import Cylinder

pytestmark = pytest.mark.gpu

TOL_K = 1e-10   # assembled K / tangent / residual entries (row-floor relative)
TOL_G = 1e-10   # exact G against the reference's central-difference G


def setup_ctx(conn, X, thick, bcs=None, temp=None):
    ctx = sg.Context(0)
    ctx.set_mesh(conn, X)
    ctx.set_material(70e9, 0.3, 2700.0, 1e-5, 10.0)
    if bcs is not None:
        ctx.set_bcs(*bcs)
    ctx.set_designs(thick, temp)
    return ctx


def cases():
    cyl = Cylinder(6, 10)
    yield "cylinder", cyl.conn, cyl.Xpts, cyl.bcs()
    conn, X = random_mesh(5, 4, 2)
    yield "distorted", conn, X, (np.array([0, 1], np.int32), np.array([0x3F, 0x3F], np.int32))


@pytest.mark.parametrize("name,conn,X,bcs", list(cases()))
def test_pattern_and_stiffness(name, conn, X, bcs):
    nn, ne = X.shape[0], conn.shape[0]
    thick = np.random.default_rng(1).uniform(0.0015, 0.0025, ne)
    ctx = setup_ctx(conn, X, thick, bcs)
    rowp, cols = ctx.pattern()
    r2, c2 = O.node_csr(nn, conn)
    assert np.array_equal(rowp, r2) and np.array_equal(cols, c2)
    colors = ctx.colors()
    for c in range(ctx.ncolors):  # no two elements of one colour share a node
        nodes = conn[colors == c].ravel()
        assert np.unique(nodes).size == nodes.size
    A = ctx.matrix()
    ctx.assemble_mattype(sg.STIFFNESS, A)
    ref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_K, 0, O.material(cte=1e-5), X, conn, thick, rowp, cols),
                          rowp, cols, *bcs)
    assert bsr_row_floor_ratio(rowp, A.values()[0], ref) < TOL_K


@pytest.mark.parametrize("name,conn,X,bcs", list(cases()))
def test_nonlinear_tangent_and_residual(name, conn, X, bcs):
    nn, ne = X.shape[0], conn.shape[0]
    rng = np.random.default_rng(2)
    thick = rng.uniform(0.0015, 0.0025, ne)
    u = 1e-3 * rng.standard_normal(6 * nn)
    ctx = setup_ctx(conn, X, thick, bcs)
    rowp, cols = ctx.pattern()
    m = O.material(cte=1e-5)
    du = ctx.vector(u)
    A = ctx.matrix()
    ctx.assemble_mattype(sg.TANGENT_NONLINEAR, A, du)
    ref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_KNL, 1, m, X, conn, thick, rowp, cols, u=u), rowp, cols, *bcs)
    assert bsr_row_floor_ratio(rowp, A.values()[0], ref) < TOL_K
    for model in (0, 1):
        R = ctx.vector()
        ctx.assemble_res(R, du, model=model)
        Rref = O.assemble_res(model, m, X, conn, thick, u=u)
        for n, msk in zip(*bcs):
            for r in range(6):
                if msk & (1 << r):
                    Rref[6 * n + r] = u[6 * n + r]
        assert np.abs(R.download() - Rref).max() <= TOL_K * np.abs(Rref).max()


@pytest.mark.parametrize("name,conn,X,bcs", list(cases()))
@pytest.mark.parametrize("temp", [0.0, 20.0])
def test_geometric_exact_and_fd(name, conn, X, bcs, temp):
    """Exact G (directional derivative) against the oracle's accurate G (the
    reference's central-difference formula at a step where it carries no
    cancellation; the tangent is quadratic in the state so the formula is
    exact for any step), and the FD replica against the reference-step FD."""
    nn, ne = X.shape[0], conn.shape[0]
    rng = np.random.default_rng(3)
    thick = rng.uniform(0.0015, 0.0025, ne)
    u = 1e-6 * rng.standard_normal(6 * nn)
    tvec = np.full(ne, temp)
    ctx = setup_ctx(conn, X, thick, bcs, temp=tvec)
    rowp, cols = ctx.pattern()
    m = O.material(cte=1e-5)
    kind = O.KIND_G if temp == 0.0 else O.KIND_G_THERMAL
    ref = O.bsr_apply_bcs(O.assemble_mat(kind, 1, m, X, conn, thick, rowp, cols, u=u, temp=tvec), rowp, cols, *bcs)
    ref_fd = O.bsr_apply_bcs(O.assemble_mat(O.KIND_G_FD, 1, m, X, conn, thick, rowp, cols, u=u, temp=tvec),
                             rowp, cols, *bcs)
    du = ctx.vector(u)
    G = ctx.matrix()
    ctx.assemble_mattype(sg.GEOMETRIC, G, du)
    Gfd = ctx.matrix()
    ctx.assemble_mattype(sg.GEOMETRIC_FD, Gfd, du)
    g, gfd = G.values()[0], Gfd.values()[0]
    e_exact = bsr_row_floor_ratio(rowp, g, ref)
    # FD against FD: both carry the reference step's cancellation noise,
    # compared relative to the largest entry of the row
    if temp == 0.0:
        e_fd = bsr_row_floor_ratio(rowp, gfd, ref_fd, floor=1.0)
    else:
        # the temperature dominates the state norm: the step along the state
        # is ~1e-4 u/|T| and the central difference keeps ~1e-7 of |G|
        e_fd = np.abs(gfd - ref_fd).max() / np.abs(ref_fd).max()
    print(f"{name} T={temp}: exact-vs-accurate {e_exact:.2e}, fd-vs-fd {e_fd:.2e}")
    assert e_exact < TOL_G
    assert e_fd < (1e-9 if temp == 0.0 else 1e-6)


def test_spmv_matches_oracle_and_batches():
    cyl = Cylinder(8, 12)
    rng = np.random.default_rng(4)
    thick = rng.uniform(0.0015, 0.0025, (3, cyl.nelems))
    ctx = setup_ctx(cyl.conn, cyl.Xpts, thick, cyl.bcs())
    rowp, cols = ctx.pattern()
    A = ctx.matrix()
    ctx.assemble_mattype(sg.STIFFNESS, A)
    vals = A.values()
    x = rng.standard_normal(3 * 6 * cyl.nnodes)
    dx, dy = ctx.vector(x), ctx.vector()
    A.mult(dx, dy)
    y = dy.download().reshape(3, -1)
    for b in range(3):
        ref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_K, 0, O.material(cte=1e-5), cyl.Xpts, cyl.conn, thick[b],
                                             rowp, cols), rowp, cols, *cyl.bcs())
        assert bsr_row_floor_ratio(rowp, vals[b], ref) < TOL_K
        yr = O.bsr_matvec(rowp, cols, vals[b], x.reshape(3, -1)[b])
        assert np.abs(y[b] - yr).max() <= 1e-13 * np.abs(yr).max()


def test_assembly_overwrites_previous_values():
    """Matrix assembly writes each block at its first touch instead of adding
    to a zeroed matrix: into a matrix holding arbitrary values, every matrix
    type (and the Jacobian entry point) gives bit for bit what it gives into
    a fresh one, two designs at a time."""
    conn, X = random_mesh(5, 4, 2)
    bcs = (np.array([0, 1], np.int32), np.array([0x3F, 0x3F], np.int32))
    rng = np.random.default_rng(17)
    thick = rng.uniform(0.0015, 0.0025, (2, conn.shape[0]))
    ctx = setup_ctx(conn, X, thick, bcs)
    u = ctx.vector(1e-3 * rng.standard_normal(2 * 6 * X.shape[0]))
    junk = 1e3 * rng.standard_normal((2, ctx.nnzb, 6, 6))
    for mattype in (sg.STIFFNESS, sg.GEOMETRIC, sg.GEOMETRIC_FD, sg.TANGENT_NONLINEAR, sg.MASS):
        fresh, used = ctx.matrix(), ctx.matrix()
        used.set_values(junk)
        ctx.assemble_mattype(mattype, fresh, u)
        ctx.assemble_mattype(mattype, used, u)
        ctx.assemble_mattype(mattype, used, u)  # twice: no accumulation
        assert np.array_equal(fresh.values(), used.values()), mattype
    for model in (sg.LINEAR, sg.NONLINEAR):
        fresh, used = ctx.matrix(), ctx.matrix()
        used.set_values(junk)
        ctx.assemble_jacobian(fresh, u=u, model=model)
        ctx.assemble_jacobian(used, u=u, model=model)
        assert np.array_equal(fresh.values(), used.values()), model
    ctx.close()
