"""shellgpu_gmres_noalloc (GMRES on the library's fixed device scratch array,
no device allocation) against shellgpu_gmres through the C ABI: the same
iterations, residual and solution bit for bit, on two calls in a row (the
second on a smaller mesh, so the array holds the first call's larger
contents), with a direct factor and without a preconditioner, for a batch of
designs, and the size limit rejected with an error."""
import numpy as np
import pytest

import shellgpu as sg
from shellgpu.problems import Cylinder

pytestmark = pytest.mark.gpu


def _ctx(cyl, nbatch):
    ctx = sg.Context(0)
    ctx.set_mesh(cyl.conn, cyl.Xpts)
    ctx.set_material(70e9, 0.3)
    ctx.set_bcs(*cyl.bcs())
    rng = np.random.default_rng(7)
    ctx.set_designs(rng.uniform(0.0015, 0.0025, (nbatch, cyl.nelems)))
    ctx.set_ordering_seed(cyl.ring_nodes(0))
    return ctx


def _compare(cyl, nbatch, precond, msub, nrestart):
    print(f"no-alloc GMRES: mesh {cyl.na}x{cyl.nc}, {nbatch} designs, precond {precond}, msub {msub}", flush=True)
    with _ctx(cyl, nbatch) as ctx:
        K = ctx.matrix()
        ctx.assemble_mattype(sg.STIFFNESS, K)
        F = None
        if precond:
            F = ctx.factor()
            F.compute(K)
        b = np.random.default_rng(3).standard_normal((nbatch, 6 * cyl.nnodes))
        db = ctx.vector(b)
        ctx.vec_zero_bcs(db)
        x1, x2 = ctx.vector(), ctx.vector()
        r1 = ctx.gmres(K, db, x1, F=F, msub=msub, nrestart=nrestart, rtol=1e-12)
        r2 = ctx.gmres_noalloc(K, db, x2, F=F, msub=msub, nrestart=nrestart, rtol=1e-12)
        a1, a2 = x1.download(), x2.download()
        assert r1 == r2
        assert np.array_equal(a1.view(np.int64), a2.view(np.int64))
        if precond:
            assert r1[1] <= 1e-12
        # the size limit: (msub + 2) vectors of this batch do not fit in 64 MiB
        big = (8 << 20) // (nbatch * 6 * cyl.nnodes) + 1
        with pytest.raises(sg.ShellGPUError, match="exceeds"):
            ctx.gmres_noalloc(K, db, x2, F=F, msub=big, nrestart=0)
        K.destroy()
        if F:
            F.destroy()
        return a1


def test_noalloc_gmres_matches_gmres_two_calls():
    # the first call on the larger mesh, the second on a smaller one
    _compare(Cylinder(20, 40), 2, True, 10, 15)
    _compare(Cylinder(8, 16), 1, True, 10, 15)


def test_noalloc_gmres_without_preconditioner():
    _compare(Cylinder(6, 12), 3, False, 30, 1)
