"""Load-stepped Newton on the nonlinear model, buckling about the nonlinear
state and its thickness gradient.

The Newton solution is checked through the oracle's nonlinear residual
(R_nl(u) - f on free dofs within 1e-9 of |f|), buckling about it against
the oracle run with the same path, and the gradient through the nonlinear
path (adjoint with the nonlinear tangent) against central differences of
the whole GPU pipeline (Newton + buckling) along seeded directions."""
import numpy as np
import pytest

import pyoracle as O
import shellgpu as sg
from shellgpu.problems import Cylinder, seeded_thickness

pytestmark = pytest.mark.gpu


def _ctx(cyl, thick):
    ctx = sg.Context(0)
    ctx.set_mesh(cyl.conn, cyl.Xpts)
    ctx.set_material(70e9, 0.3)
    ctx.set_bcs(*cyl.bcs())
    ctx.set_designs(thick)
    ctx.set_ordering_seed(cyl.ring_nodes(0))
    return ctx


def _free(cyl, v):
    v = v.copy()
    bn, bm = cyl.bcs()
    for node, mask in zip(bn, bm):
        for k in range(6):
            if mask >> k & 1:
                v[6 * node + k] = 0.0
    return v


def test_newton_solves_oracle_residual():
    cyl = Cylinder(10, 20, imperfection=(0.0002, 2, 3))
    thick = seeded_thickness(cyl.nelems, 71)
    load = 0.3 * cyl.classical_load(0.002, 70e9, 0.3)
    f = cyl.edge_load(load)
    ctx = _ctx(cyl, thick)
    u, its, rel = ctx.newton_solve(f, nsteps=5, max_iters=20, rtol=1e-10)
    u = u.download()
    print("Newton iterations", its, "relres", rel)
    assert rel[0] <= 1e-10 and its >= 5
    r = _free(cyl, O.assemble_res(1, O.material(), cyl.Xpts, cyl.conn, thick[0], u=u) - f)
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(_free(cyl, f))
    # the load is large enough for the path to be visibly nonlinear
    rl = _free(cyl, O.assemble_res(0, O.material(), cyl.Xpts, cyl.conn, thick[0], u=u) - f)
    assert np.linalg.norm(rl) >= 1e-4 * np.linalg.norm(_free(cyl, f))


def test_buckling_about_nonlinear_state_and_gradient():
    cyl = Cylinder(10, 20, imperfection=(0.0002, 2, 3))
    thick = seeded_thickness(cyl.nelems, 72)
    load = 0.2 * cyl.classical_load(0.002, 70e9, 0.3)
    f = cyl.edge_load(load)
    nev = 3

    def run(th, grad=False):
        ctx = _ctx(cyl, th)
        u, _, rel = ctx.newton_solve(f, nsteps=5, max_iters=20, rtol=1e-12)
        B = ctx.buckling(sigma=0.0, nev=nev, max_lanczos=80, tol=1e-12, model=sg.NONLINEAR)
        B.solve(u0=u)
        lam = B.eigenvalues()[0][0]
        sig = B.info()["sigma"][0]
        g = B.eigen_sens(np.eye(nev)[0])[0][0] if grad else None
        uh = u.download()
        B.destroy()
        ctx.close()
        return lam, g, sig, uh

    lam, g, sig, u = run(thick, grad=True)
    # oracle buckling about the same path
    lo = O.linear_buckling(cyl.Xpts, cyl.conn, thick[0], cyl.bcs(), f, sig, nev=nev, max_lanczos=80, u0=u)[0]
    print("gpu", lam, "oracle", lo)
    assert np.abs(lam - lo).max() <= 1e-8 * np.abs(lo).max()
    rng = np.random.default_rng(3002)
    h = 1e-4
    for k in range(2):
        dirn = rng.standard_normal(cyl.nelems) * thick[0]
        lp = run(thick + h * dirn)[0]
        lm = run(thick - h * dirn)[0]
        fd = (lp[0] - lm[0]) / (2 * h)
        print(k, g @ dirn, fd)
        assert abs(g @ dirn - fd) <= 1e-6 * abs(lam[0]), k


def test_shared_newton_factor_matches_own_factors():
    # Newton from the zero state factors its tangent at zero state (= K) into
    # the caller's factor; the buckling analysis then uses it for the shift
    # estimate and the path adjoint instead of factoring K itself.  Loads,
    # shift-independent, and the gradient agree with the unshared run.
    cyl = Cylinder(10, 20, imperfection=(0.0002, 2, 3))
    thick = seeded_thickness(cyl.nelems, 73)
    f = cyl.edge_load(0.2 * cyl.classical_load(0.002, 70e9, 0.3))
    nev = 3
    out = []
    for shared in (False, True):
        ctx = _ctx(cyl, thick)
        F = ctx.factor() if shared else None
        u, _, rel = ctx.newton_solve(f, nsteps=5, max_iters=20, rtol=1e-12, factor=F)
        B = ctx.buckling(sigma=0.0, nev=nev, max_lanczos=80, tol=1e-12, model=sg.NONLINEAR)
        if shared:
            B.set_stiffness_factor(F)
        B.solve(u0=u)
        lam, err = B.eigenvalues()
        g = B.eigen_sens(np.eye(nev)[0])[0][0]
        out.append((lam[0], g, u.download()))
        B.destroy()
        if F is not None:
            F.destroy()
        ctx.close()
    (l0, g0, u0), (l1, g1, u1) = out
    print("own", l0, "shared", l1)
    assert np.abs(u1 - u0).max() <= 1e-12 * np.abs(u0).max()
    assert np.abs(l1 - l0).max() <= 1e-10 * np.abs(l0).max()
    assert np.abs(g1 - g0).max() <= 1e-8 * np.abs(g0).max()
    # a factor that holds no factorisation is refused
    ctx = _ctx(cyl, thick)
    B = ctx.buckling(sigma=0.0, nev=nev, max_lanczos=80, tol=1e-12, model=sg.NONLINEAR)
    F = ctx.factor()
    with pytest.raises(sg.ShellGPUError):
        B.set_stiffness_factor(F)  # holds no factorisation yet
    B.destroy()
    F.destroy()
    ctx.close()
