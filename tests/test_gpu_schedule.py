"""The autotuner's schedules: every compiled schedule of the matrix-assembly
kernels (BSR block reads before the quadrature loop or at the scatter, one-
or two-wavefront register budget, two elements per workgroup or wide
workgroups of 4 with the geometry spread over them, on colours of
ragged length) and both element colourings assemble the same matrices and
residuals (the same arithmetic; within round-off of the default schedule,
which the oracle tests pin), and shellgpu_tune reports and keeps valid
choices, and that the timing counters count every colour launch."""
import numpy as np
import pytest

import shellgpu as sg
from helpers import random_mesh
from shellgpu.problems import Cylinder

pytestmark = pytest.mark.gpu


def _ctx(conn, X, thick, bcs):
    ctx = sg.Context(0)
    ctx.set_mesh(conn, X)
    ctx.set_material(70e9, 0.3, 2700.0, 0.0, 10.0)
    ctx.set_bcs(*bcs)
    ctx.set_designs(thick)
    return ctx


def _assemble_all(ctx, u, nn):
    out = {}
    for name, mt in (("K", sg.STIFFNESS), ("Kt", sg.TANGENT_NONLINEAR), ("G", sg.GEOMETRIC)):
        A = ctx.matrix()
        ctx.assemble_mattype(mt, A, u)
        out[name] = A.values()
        A.destroy()
    for model in (sg.LINEAR, sg.NONLINEAR):
        A, R = ctx.matrix(), ctx.vector()
        ctx.assemble_jacobian(A, R, u, model=model, gamma=1e6)
        out[f"J{model}"], out[f"R{model}"] = A.values(), R.download()
        A.destroy()
    return out


def _close(a, b):
    return np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_every_schedule_assembles_the_same_matrices():
    conn, X = random_mesh(6, 5, seed=81)
    rng = np.random.default_rng(82)
    thick = rng.uniform(0.0015, 0.0025, (2, conn.shape[0]))
    bcs = (np.array([0, 1], np.int32), np.array([0x3F, 0x3F], np.int32))
    with _ctx(conn, X, thick, bcs) as ctx:
        u = ctx.vector(1e-3 * rng.standard_normal(2 * 6 * X.shape[0]))
        ref = _assemble_all(ctx, u, X.shape[0])
        for v in range(1, 8):
            for c in range(5):
                ctx.set_elem_schedule(c, v)
            got = _assemble_all(ctx, u, X.shape[0])
            for k in ref:
                assert _close(got[k], ref[k]), (v, k)


def test_tune_keeps_valid_choices_and_results():
    cyl = Cylinder(12, 24)
    rng = np.random.default_rng(83)
    thick = rng.uniform(0.0015, 0.0025, (1, cyl.nelems))
    with _ctx(cyl.conn, cyl.Xpts, thick, cyl.bcs()) as ctx:
        u = ctx.vector(1e-4 * rng.standard_normal(6 * cyl.nnodes))
        ref = _assemble_all(ctx, u, cyl.nnodes)
        rep = ctx.tune(coloring=True)
        assert set(rep["elem"]) == {"stiffness", "tangent", "geometric", "jacobian_linear", "jacobian_nonlinear"}
        assert all(0 <= r["schedule"] < 8 and len(r["us"]) == 8 for r in rep["elem"].values())
        assert rep["coloring"]["colors"] == ctx.ncolors
        # the colouring in force: no two elements of a colour share a node
        col = ctx.colors()
        for c in range(ctx.ncolors):
            nodes = cyl.conn[col == c].ravel()
            assert np.unique(nodes).size == nodes.size
        got = _assemble_all(ctx, u, cyl.nnodes)
        for k in ref:
            assert _close(got[k], ref[k]), k


def test_most_neighbours_first_coloring_is_valid_on_a_junction_mesh():
    from shellgpu.problems import BoxBeam
    b = BoxBeam(20, 8, 2, 4, span=1.0, chord=0.2, height=0.05)
    thick = np.full((1, b.nelems), 0.004)
    with _ctx(b.conn, b.Xpts, thick, b.bcs()) as ctx:
        u = ctx.vector(1e-5 * np.random.default_rng(84).standard_normal(6 * b.nnodes))
        ref = _assemble_all(ctx, u, b.nnodes)
        ctx.tune(coloring=True)
        col = ctx.colors()
        for c in range(ctx.ncolors):
            nodes = b.conn[col == c].ravel()
            assert np.unique(nodes).size == nodes.size
        got = _assemble_all(ctx, u, b.nnodes)
        for k in ref:
            assert _close(got[k], ref[k]), k


@pytest.mark.parametrize("nx, ny", [(1, 1), (3, 1)])
def test_schedules_on_colours_shorter_than_a_workgroup(nx, ny):
    """Colours of one or two elements: the wide schedules' padding slots
    compute a copy of the colour's last element and write nothing, so every
    schedule assembles what the default one does."""
    conn, X = random_mesh(nx, ny, seed=84)
    rng = np.random.default_rng(85)
    thick = rng.uniform(0.0015, 0.0025, (1, conn.shape[0]))
    bcs = (np.array([0], np.int32), np.array([0x3F], np.int32))
    with _ctx(conn, X, thick, bcs) as ctx:
        assert ctx.ncolors == (1 if conn.shape[0] == 1 else 2)
        u = ctx.vector(1e-3 * rng.standard_normal(6 * X.shape[0]))
        ref = _assemble_all(ctx, u, X.shape[0])
        for v in range(1, 8):
            for c in range(5):
                ctx.set_elem_schedule(c, v)
            got = _assemble_all(ctx, u, X.shape[0])
            for k in ref:
                assert _close(got[k], ref[k]), (v, k)


def test_timing_counts_colour_launches():
    """With timing on, each assembly pass is one event pair around its colour
    launches; get_timing reports every colour launch and the SpMVs, then
    resets."""
    cyl = Cylinder(8, 16)
    thick = np.full((1, cyl.nelems), 0.002)
    with _ctx(cyl.conn, cyl.Xpts, thick, cyl.bcs()) as ctx:
        A = ctx.matrix()
        x, y = ctx.vector(np.ones(6 * cyl.nnodes)), ctx.vector()
        ctx.set_timing(True)
        for _ in range(3):
            ctx.assemble_mattype(sg.STIFFNESS, A)
        A.mult(x, y)
        elem_ms, elem_n, spmv_ms, spmv_n = ctx.get_timing()
        assert elem_n == 3 * ctx.ncolors and elem_ms > 0.0
        assert spmv_n == 1 and spmv_ms > 0.0
        assert ctx.get_timing()[1] == 0
        ctx.set_timing(False)
        ctx.assemble_mattype(sg.STIFFNESS, A)
        assert ctx.get_timing()[1] == 0
        A.destroy()
