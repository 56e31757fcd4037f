"""The benchmark configurations B-E at their full sizes on one GPU (the
20 x 40 cylinder of configuration A is covered by test_gpu_solver.py).

B  200 x 400 cylinder (80,000 quads, 482,400 dof) with seeded thickness
   (seed 1234) and radial imperfection (seed 1235), unit axial edge load:
   5-step Newton on the nonlinear model, checked through the oracle's
   nonlinear residual; buckling about the final state, first 5 loads, run
   with both factorisations (cyclic reduction and the sequential form), which
   agree within 1e-10 (loads) and 1e-6 rad (principal angles of the modes,
   K inner product); every load within 1e-8 of the Rayleigh quotient of its
   mode with the oracle's K and G(u); K and G(u) themselves against the
   oracle's assembly (row-floored relative error 1e-10; G from the oracle's
   long-double evaluation, whose double evaluation carries ~1e-10 of
   round-off on this mesh).
C  64 thickness designs (seeds 2000..2063) of the 100 x 200 cylinder in one
   batch: every design's 5 loads (1e-10) and KS thickness gradient (1e-8)
   against its own single-design run; K and G(path) of two designs of the
   batch against the oracle (1e-10).
D  the KS (rho = 100) aggregate of the first 5 loads of the 200 x 400
   cylinder (linear static path): its gradient with respect to all 80,000
   thicknesses and 241,200 coordinates against central differences of the
   GPU's loads at relative step 1e-6 along 20 seeded directions (seed 3000):
   10 thickness directions t * N(0, 1) and 10 coordinate directions
   (L / 200) N(0, 1) per coordinate.
E  the 128,000-quad stiffened box beam, 16 load cases in one batch: every
   case's static path through the oracle's K (normwise backward error
   1e-11),
   and the panel-thickness gradient of the summed KS of the positive loads
   against central differences along 2 seeded panel directions at relative
   steps 1e-4 and 1e-6 (the bound adds the loads' round-off over the step).

Every test prints what it runs before each GPU step.
"""
import time

import numpy as np
import pytest

import pyoracle as O
import shellgpu as sg
from helpers import bsr_row_floor_ratio
from shellgpu.problems import BoxBeam, Cylinder, seeded_imperfection, seeded_thickness

pytestmark = pytest.mark.gpu

KS_RHO = 100.0
# relative step of the central differences of configuration D
FD_STEP = 1e-6


def log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def cyl_ctx(cyl, thick, Xpts=None):
    ctx = sg.Context(0)
    ctx.set_mesh(cyl.conn, cyl.Xpts if Xpts is None else Xpts)
    ctx.set_material(70e9, 0.3)
    ctx.set_bcs(*cyl.bcs())
    ctx.set_designs(thick)
    ctx.set_ordering_seed(cyl.ring_nodes(0))
    return ctx


def bc_dofs(nnodes, bcs):
    """Boolean mask of the BC dofs (6 per node)."""
    m = np.zeros((nnodes, 6), dtype=bool)
    nodes, mask = bcs[0], bcs[1]
    for k in range(6):
        m[nodes[(mask >> k) & 1 == 1], k] = True
    return m.ravel()


def ks_of_loads(lam):
    """KS (rho = 100) of -lambda_i / lambda_1 (a smooth minimum of the loads
    normalised by the first, the normalisation differentiated too) per row:
    values and d KS / d lambda."""
    ks, coef = [], []
    for row in np.atleast_2d(lam):
        v, c = sg.ks_aggregate(-row / row[0], KS_RHO)
        d = -c / row[0]
        d[0] += float(c @ row) / row[0] ** 2
        ks.append(v)
        coef.append(d)
    return np.array(ks), np.array(coef)


def clusters(lam, rel=1e-6):
    """Index groups of (nearly) equal loads."""
    groups, cur = [], [0]
    for j in range(1, len(lam)):
        if abs(lam[j] - lam[cur[-1]]) <= rel * abs(lam[j]):
            cur.append(j)
        else:
            groups.append(cur)
            cur = [j]
    groups.append(cur)
    return groups


def k_angle(V, W, rowp, cols, Kvals):
    """Largest principal angle between span(V) and span(W) (rows K-orthonormal)."""
    KW = np.stack([O.bsr_matvec(rowp, cols, Kvals, w) for w in W])
    s = np.linalg.svd(V @ KW.T, compute_uv=False)
    return float(np.arccos(min(1.0, s.min())))


# ---------------------------------------------------------------------------
# configuration B
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cyl_b():
    cyl = Cylinder(200, 400, R=1.0, L=2.0, imperfection=seeded_imperfection(1235))
    thick = seeded_thickness(cyl.nelems, 1234)
    return cyl, thick


def test_config_b_newton_and_buckling(cyl_b, monkeypatch):
    cyl, thick = cyl_b
    f = cyl.edge_load(1.0)
    bc = bc_dofs(cyl.nnodes, cyl.bcs())
    m = O.material()
    runs = {}
    for factor in ("cr", "thomas"):
        monkeypatch.setenv("SHELLGPU_FACTOR", factor)
        log(f"config B (200x400): Newton + buckling, factorisation {factor}")
        with cyl_ctx(cyl, thick) as ctx:
            du, its, rel = ctx.newton_solve(f, nsteps=5, max_iters=20, rtol=1e-10)
            B = ctx.buckling(sigma=0.0, nev=5, max_lanczos=150, tol=1e-10, model=sg.NONLINEAR)
            B.solve(u0=du)
            lam, err = B.eigenvalues()
            phi = np.stack([B.eigenvector(j)[0] for j in range(5)])
            u = du.download()
            log(f"  Newton {its} iterations, relres {rel[0]:.2e}; loads {lam[0]}; Ritz residuals {err[0]}")
            assert rel[0] <= 1e-10
            assert err[0].max() <= 1e-10
            out = {"lam": lam[0], "phi": phi, "u": u}
            if factor == "cr":
                K, G = ctx.matrix(), ctx.matrix()
                ctx.assemble_mattype(sg.STIFFNESS, K)
                ctx.assemble_mattype(sg.GEOMETRIC, G, du)
                out["K"], out["G"] = K.values()[0], G.values()[0]
                K.destroy()
                G.destroy()
                rowp, cols = ctx.pattern()
            B.destroy()
        runs[factor] = out

    cr, th = runs["cr"], runs["thomas"]
    log("config B: oracle residual, K, G(u), Rayleigh quotients")
    # the Newton state solves the oracle's nonlinear residual
    r = O.assemble_res(1, m, cyl.Xpts, cyl.conn, thick[0], u=cr["u"]) - f
    r[bc] = 0.0
    fr = f.copy()
    fr[bc] = 0.0
    log(f"  |R_nl(u) - f| / |f| = {np.linalg.norm(r) / np.linalg.norm(fr):.2e}")
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(fr)
    # K and G(u) against the oracle's assembly at full size
    Kref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_K, 0, m, cyl.Xpts, cyl.conn, thick[0], rowp, cols), rowp, cols,
                           *cyl.bcs())
    ek = bsr_row_floor_ratio(rowp, cr["K"], Kref)
    # the geometric stiffness reference in long double: the double central
    # difference carries ~1e-10 of round-off on this mesh
    Gref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_G_LD, 1, m, cyl.Xpts, cyl.conn, thick[0], rowp, cols,
                                          u=cr["u"]), rowp, cols, *cyl.bcs())
    eg = bsr_row_floor_ratio(rowp, cr["G"], Gref)
    log(f"  K parity {ek:.2e}, G(u) parity {eg:.2e}")
    assert ek <= 1e-10 and eg <= 1e-10
    # each load is the Rayleigh quotient of its mode with the oracle's matrices
    for j in range(5):
        p = cr["phi"][j]
        rq = -(p @ O.bsr_matvec(rowp, cols, Kref, p)) / (p @ O.bsr_matvec(rowp, cols, Gref, p))
        log(f"  load {j}: {cr['lam'][j]:.12e}, Rayleigh quotient {rq:.12e}")
        assert abs(rq - cr["lam"][j]) <= 1e-8 * abs(cr["lam"][j])
    # both factorisations: the same loads and modes
    dl = np.abs(cr["lam"] - th["lam"]).max() / np.abs(cr["lam"]).max()
    log(f"  cyclic reduction vs sequential factor: loads {dl:.2e}")
    assert dl <= 1e-10
    groups = clusters(cr["lam"])
    if len(groups) > 1:
        groups = groups[:-1]  # the last cluster may continue past the 5th load
    for g in groups:
        a = k_angle(cr["phi"][g], th["phi"][g], rowp, cols, Kref)
        log(f"  modes {g}: principal angle {a:.2e} rad")
        assert a <= 1e-6


# ---------------------------------------------------------------------------
# configuration D
# ---------------------------------------------------------------------------
def d_analysis(cyl, thick, f, Xpts=None, sigma=0.0, grad=False):
    """Linear static path, buckling (first 5 loads), KS of the loads; the
    gradient with respect to thickness and coordinates when grad."""
    with cyl_ctx(cyl, thick, Xpts) as ctx:
        B = ctx.buckling(sigma=sigma, nev=5, max_lanczos=150, tol=1e-12)
        B.solve(force=f)
        lam, err = B.eigenvalues()
        assert err.max() <= 1e-12
        ks, coef = ks_of_loads(lam)
        out = {"lam": lam[0], "ks": ks[0], "sigma": B.info()["sigma"][0]}
        if grad:
            gt, gX = B.eigen_sens(coef, thickness=True, coords=True)
            out["gt"], out["gX"] = gt[0], gX[0]
        B.destroy()
    return out


@pytest.fixture(scope="module")
def grad_d(cyl_b):
    cyl, thick = cyl_b
    log("config D (200x400): buckling and KS gradient (thickness and coordinates)")
    return d_analysis(cyl, thick, cyl.edge_load(1.0), grad=True)


def d_direction(kind, k, cyl, thick):
    """Direction k of kind 't' or 'X' (seed 3000 + ...)."""
    rng = np.random.default_rng(3000 + (0 if kind == "t" else 100) + k)
    if kind == "t":
        return thick[0] * rng.standard_normal(cyl.nelems)
    return (cyl.L / cyl.na) * rng.standard_normal(3 * cyl.nnodes)


@pytest.mark.parametrize("kind,first", [("t", 0), ("t", 5), ("X", 0), ("X", 5)])
def test_config_d_ks_gradient_central_differences(cyl_b, grad_d, kind, first):
    cyl, thick = cyl_b
    f = cyl.edge_load(1.0)
    sig = grad_d["sigma"]
    g = grad_d["gt"] if kind == "t" else grad_d["gX"]
    for k in range(first, first + 5):
        v = d_direction(kind, k, cyl, thick)
        vals = []
        for s in (1.0, -1.0):
            log(f"config D: direction {kind}{k}, step {s * FD_STEP:+.0e}")
            if kind == "t":
                vals.append(d_analysis(cyl, thick + s * FD_STEP * v, f, sigma=sig)["ks"])
            else:
                X = (cyl.Xpts.ravel() + s * FD_STEP * v).reshape(-1, 3)
                vals.append(d_analysis(cyl, thick, f, Xpts=X, sigma=sig)["ks"])
        fd = (vals[0] - vals[1]) / (2 * FD_STEP)
        an = float(g @ v)
        # the KS values carry round-off of about 1e-13 (loads converged to
        # Ritz residuals of 1e-12, assembled and factored in FP64), which the
        # difference quotient amplifies to ~1e-13 / (2 FD_STEP) = 5e-8
        log(f"  adjoint {an:.10e}, central difference {fd:.10e}, rel {abs(an - fd) / abs(an):.2e}")
        assert abs(an - fd) <= 1e-5 * abs(an) + 5e-7, (kind, k)


# ---------------------------------------------------------------------------
# configuration C
# ---------------------------------------------------------------------------
C_DESIGNS = 64
C_CHECK = (0, 41)  # designs whose K and G are compared with the oracle


@pytest.fixture(scope="module")
def batch_c():
    cyl = Cylinder(100, 200, R=1.0, L=2.0)
    thick = np.concatenate([seeded_thickness(cyl.nelems, 2000 + d) for d in range(C_DESIGNS)])
    f = cyl.edge_load(1.0)
    log(f"config C: {C_DESIGNS} designs of the 100x200 cylinder in one batch")
    with cyl_ctx(cyl, thick) as ctx:
        B = ctx.buckling(sigma=0.0, nev=5, max_lanczos=120, tol=1e-12)
        B.solve(force=f)
        lam, err = B.eigenvalues()
        assert err.max() <= 1e-12
        _, coef = ks_of_loads(lam)
        gt = B.eigen_sens(coef)[0]
        path = B.info()["path"]
        B.destroy()
        log("config C: K and G(path) of the batch")
        K, G = ctx.matrix(), ctx.matrix()
        ctx.assemble_mattype(sg.STIFFNESS, K)
        ctx.assemble_mattype(sg.GEOMETRIC, G, ctx.vector(path))
        mats = {d: (K.design_values(d), G.design_values(d)) for d in C_CHECK}
        K.destroy()
        G.destroy()
        rowp, cols = ctx.pattern()
    return {"cyl": cyl, "thick": thick, "f": f, "lam": lam, "gt": gt, "path": path, "mats": mats,
            "rowp": rowp, "cols": cols}


def test_config_c_matrices_match_oracle(batch_c):
    cyl, thick, rowp, cols = batch_c["cyl"], batch_c["thick"], batch_c["rowp"], batch_c["cols"]
    m = O.material()
    for d in C_CHECK:
        Kd, Gd = batch_c["mats"][d]
        Kref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_K, 0, m, cyl.Xpts, cyl.conn, thick[d], rowp, cols), rowp, cols,
                               *cyl.bcs())
        Gref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_G_LD, 1, m, cyl.Xpts, cyl.conn, thick[d], rowp, cols,
                                              u=batch_c["path"][d]), rowp, cols, *cyl.bcs())
        ek, eg = bsr_row_floor_ratio(rowp, Kd, Kref), bsr_row_floor_ratio(rowp, Gd, Gref)
        log(f"config C design {d}: K parity {ek:.2e}, G parity {eg:.2e}")
        assert ek <= 1e-10 and eg <= 1e-10


@pytest.mark.parametrize("first", [0, 16, 32, 48])
def test_config_c_batch_matches_single_designs(batch_c, first):
    cyl, thick, f = batch_c["cyl"], batch_c["thick"], batch_c["f"]
    for d in range(first, first + 16):
        log(f"config C: design {d} alone")
        with cyl_ctx(cyl, thick[d:d + 1]) as ctx:
            B = ctx.buckling(sigma=0.0, nev=5, max_lanczos=120, tol=1e-12)
            B.solve(force=f)
            lam = B.eigenvalues()[0]
            _, coef = ks_of_loads(lam)
            g1 = B.eigen_sens(coef)[0][0]
            B.destroy()
        dl = np.abs(lam[0] - batch_c["lam"][d]).max() / np.abs(lam[0]).max()
        dg = np.abs(g1 - batch_c["gt"][d]).max() / np.abs(g1).max()
        assert dl <= 1e-10, (d, dl)
        assert dg <= 1e-8, (d, dg)


# ---------------------------------------------------------------------------
# configuration E
# ---------------------------------------------------------------------------
def positive_ks(lam, npos=5):
    """Per load case: KS of the first npos positive loads and its
    coefficients on the full eigenvalue list."""
    ks, coef = [], np.zeros_like(lam)
    for d, row in enumerate(lam):
        idx = np.where(row > 0)[0][:npos]
        assert idx.size == npos
        v, c = ks_of_loads(row[idx])
        ks.append(v[0])
        coef[d, idx] = c[0]
    return np.array(ks), coef


def box_analysis(b, t, forces, grad=False):
    with sg.Context(0) as ctx:
        ctx.set_mesh(b.conn, b.Xpts)
        ctx.set_material(70e9, 0.3)
        ctx.set_bcs(*b.bcs())
        ctx.set_designs(np.tile(t, (len(forces), 1)))
        ctx.set_ordering_seed(b.root_nodes())
        B = ctx.buckling(sigma=0.0, nev=8, max_lanczos=150, tol=1e-12)
        B.solve(force=forces)
        lam, err = B.eigenvalues()
        ks, coef = positive_ks(lam)
        out = {"lam": lam, "err": err, "ks": ks}
        if grad:
            out["gt"] = B.eigen_sens(coef)[0]
            out["path"] = B.info()["path"]
            out["rowp"], out["cols"] = ctx.pattern()
        B.destroy()
    return out


def test_config_e_box_beam():
    b = BoxBeam()
    assert b.nelems == 128000
    t = b.panel_thickness(4000)
    forces = np.stack([b.load_case(c, 16) for c in range(16)])
    log(f"config E: box beam ({b.nelems} quads, {6 * b.nnodes} dof), 16 load cases")
    base = box_analysis(b, t, forces, grad=True)
    assert base["err"][base["lam"] > 0].max() <= 1e-10
    log("config E: static paths through the oracle's K")
    rowp, cols = base["rowp"], base["cols"]
    Kref = O.bsr_apply_bcs(O.assemble_mat(O.KIND_K, 0, O.material(), b.Xpts, b.conn, t, rowp, cols), rowp, cols,
                           *b.bcs())
    bc = bc_dofs(b.nnodes, b.bcs())
    Kabs = np.abs(Kref)
    for c in range(16):
        u = base["path"][c]
        r = O.bsr_matvec(rowp, cols, Kref, u) - forces[c]
        r[bc] = 0.0
        # normwise backward error |K u - f| / | |K| |u| + |f| |: the GPU's K
        # and the oracle's differ by round-off, and |K| |u| exceeds |f| by
        # orders of magnitude on this structure (the relative residual
        # |K u - f| / |f| is 1e-8 here while the GPU solve reached 1e-12)
        scale = O.bsr_matvec(rowp, cols, Kabs, np.abs(u)) + np.abs(forces[c])
        scale[bc] = 0.0
        omega = np.linalg.norm(r) / np.linalg.norm(scale)
        log(f"  case {c}: |Ku - f| / |f| = {np.linalg.norm(r) / np.linalg.norm(forces[c]):.2e}, "
            f"backward error {omega:.2e}")
        assert omega <= 1e-11, (c, omega)
    # panel-thickness gradient of sum_c KS_c
    gpan = np.bincount(b.panel_of, weights=base["gt"].sum(axis=0), minlength=b.npanels)
    rng = np.random.default_rng(4002)
    for k in range(2):
        vp = rng.standard_normal(b.npanels) * np.bincount(b.panel_of, weights=t, minlength=b.npanels) / \
            np.bincount(b.panel_of, minlength=b.npanels)
        an = float(gpan @ vp)
        for h in (1e-4, FD_STEP):
            vals = []
            for s in (1.0, -1.0):
                log(f"config E: panel direction {k}, step {s * h:+.0e}")
                vals.append(box_analysis(b, t + s * h * vp[b.panel_of], forces)["ks"].sum())
            fd = (vals[0] - vals[1]) / (2 * h)
            log(f"  adjoint {an:.10e}, central difference {fd:.10e}, rel {abs(an - fd) / abs(an):.2e}")
            # the loads of this structure carry ~1e-12 relative round-off
            # (factors of a stiffness of high condition), so each KS value
            # (of order 1, 16 of them) carries ~1e-12 and the quotient
            # 16e-12 / h; the truncation error is O(h^2)
            assert abs(an - fd) <= 1e-6 * abs(an) + 16 * 1e-12 / h, (k, h)
