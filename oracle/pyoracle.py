"""ctypes binding of oracle/liboracle.so -- the CPU FP64 restatement of the
reference path.  TEST INFRASTRUCTURE: imported only by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


class Material(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("rho", C.c_double),
                ("cte", C.c_double), ("k_drill", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        _lib = C.CDLL(LIB)
        _lib.or_elem_energy.restype = C.c_double
    return _lib


def _d(a):
    return a.ctypes.data_as(_DP)


def _i(a):
    return a.ctypes.data_as(_IP)


def material(E=70e9, nu=0.3, rho=2700.0, cte=0.0, k_drill=10.0):
    return Material(E, nu, rho, cte, k_drill)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def elem_residual(model, mat, t, temp, X, q):
    res = np.zeros(24)
    lib().or_elem_residual(model, C.byref(mat), C.c_double(t), C.c_double(temp), _d(_f64(X)), _d(_f64(q)), _d(res))
    return res


def elem_stiffness(model, mat, t, temp, X, q):
    K = np.zeros(576)
    lib().or_elem_stiffness(model, C.byref(mat), C.c_double(t), C.c_double(temp), _d(_f64(X)), _d(_f64(q)), _d(K))
    return K.reshape(24, 24)


def elem_output(model, mat, t, ys, X, q):
    out = np.zeros(124)
    lib().or_elem_output(model, C.byref(mat), C.c_double(t), C.c_double(ys), _d(_f64(X)), _d(_f64(q)), _d(out))
    return out.reshape(4, 31)


def elem_mass(mat, t, X):
    M = np.zeros(576)
    lib().or_elem_mass(C.byref(mat), C.c_double(t), _d(_f64(X)), _d(M))
    return M.reshape(24, 24)


def elem_jacobian(model, mat, t, temp, X, q, alpha=1.0):
    K = np.zeros(576)
    res = np.zeros(24)
    lib().or_elem_jacobian(model, C.byref(mat), C.c_double(t), C.c_double(temp), C.c_double(alpha),
                           _d(_f64(X)), _d(_f64(q)), _d(res), _d(K))
    return res, K.reshape(24, 24)


def elem_geometric_step_ld(mat, t, temp, X, u, dh=1.0):
    G = np.zeros(576)
    lib().or_elem_geometric_step_ld(C.byref(mat), C.c_double(t), C.c_double(temp), _d(_f64(X)), _d(_f64(u)),
                                    C.c_double(dh), _d(G))
    return G.reshape(24, 24)


def elem_geometric_fd(mat, t, temp, X, u):
    G = np.zeros(576)
    lib().or_elem_geometric_fd(C.byref(mat), C.c_double(t), C.c_double(temp), _d(_f64(X)), _d(_f64(u)), _d(G))
    return G.reshape(24, 24)


def elem_geometric_step(mat, t, temp, X, u, dh):
    G = np.zeros(576)
    lib().or_elem_geometric_step(C.byref(mat), C.c_double(t), C.c_double(temp), _d(_f64(X)), _d(_f64(u)),
                                 C.c_double(dh), _d(G))
    return G.reshape(24, 24)


def elem_adj_res_dv(model, mat, t, temp, psi, X, q, scale=1.0):
    out = C.c_double(0.0)
    lib().or_elem_adj_res_dv(model, C.byref(mat), C.c_double(t), C.c_double(temp), C.c_double(scale),
                             _d(_f64(psi)), _d(_f64(X)), _d(_f64(q)), C.byref(out))
    return out.value


def elem_energy(model, mat, t, temp, X, q):
    return lib().or_elem_energy(model, C.byref(mat), C.c_double(t), C.c_double(temp), _d(_f64(X)), _d(_f64(q)))


def node_csr(nnodes, conn):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    rp, cp = _IP(), _IP()
    nnz = lib().or_node_csr(nnodes, conn.size // 4, _i(conn), C.byref(rp), C.byref(cp))
    rowp = np.ctypeslib.as_array(rp, shape=(nnodes + 1,)).copy()
    cols = np.ctypeslib.as_array(cp, shape=(nnz,)).copy()
    libc = C.CDLL(None)
    libc.free(C.cast(rp, C.c_void_p))
    libc.free(C.cast(cp, C.c_void_p))
    return rowp, cols


def first_appearance(nnodes, conn):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    out = np.empty(nnodes, np.int32)
    lib().or_first_appearance(nnodes, conn.size // 4, _i(conn), _i(out))
    return out


KIND_K, KIND_G_FD, KIND_KNL, KIND_G, KIND_G_THERMAL, KIND_M = 0, 1, 2, 3, 4, 5
# the accurate geometric stiffness (KIND_G's formula) evaluated in long double
KIND_G_LD = 6


def assemble_mat(kind, model, mat, Xpts, conn, thick, rowp, cols, u=None, temp=None):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    nnodes = Xpts.size // 3
    nelems = conn.size // 4
    vals = np.zeros(36 * cols.size)
    lib().or_assemble_mat(kind, model, C.byref(mat), nnodes, nelems, _i(conn), _d(_f64(Xpts)), _d(_f64(thick)),
                          _d(_f64(temp)) if temp is not None else None, _d(_f64(u)) if u is not None else None,
                          _i(np.ascontiguousarray(rowp, np.int32)), _i(np.ascontiguousarray(cols, np.int32)),
                          _d(vals))
    return vals.reshape(-1, 6, 6)


def assemble_res(model, mat, Xpts, conn, thick, u=None, temp=None):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    nnodes = Xpts.size // 3
    R = np.zeros(6 * nnodes)
    lib().or_assemble_res(model, C.byref(mat), nnodes, conn.size // 4, _i(conn), _d(_f64(Xpts)), _d(_f64(thick)),
                          _d(_f64(temp)) if temp is not None else None, _d(_f64(u)) if u is not None else None, _d(R))
    return R


def bsr_apply_bcs(vals, rowp, cols, bc_nodes, bc_mask, diag=1.0):
    vals = np.ascontiguousarray(vals, dtype=np.float64).copy()
    lib().or_bsr_apply_bcs(rowp.size - 1, _i(np.ascontiguousarray(rowp, np.int32)),
                           _i(np.ascontiguousarray(cols, np.int32)), _d(vals), len(bc_nodes),
                           _i(np.ascontiguousarray(bc_nodes, np.int32)), _i(np.ascontiguousarray(bc_mask, np.int32)),
                           C.c_double(diag))
    return vals


def bsr_matvec(rowp, cols, vals, x):
    y = np.zeros(6 * (rowp.size - 1))
    lib().or_bsr_matvec(rowp.size - 1, _i(np.ascontiguousarray(rowp, np.int32)),
                        _i(np.ascontiguousarray(cols, np.int32)), _d(_f64(vals)), _d(_f64(x)), _d(y))
    return y


def bsr_to_dense(rowp, cols, vals):
    n = rowp.size - 1
    A = np.zeros((6 * n, 6 * n))
    vals = np.asarray(vals).reshape(-1, 6, 6)
    for i in range(n):
        for p in range(rowp[i], rowp[i + 1]):
            A[6 * i:6 * i + 6, 6 * cols[p]:6 * cols[p] + 6] = vals[p]
    return A


class BucklingProblem(C.Structure):
    _fields_ = [("nnodes", C.c_int), ("nelems", C.c_int), ("conn", _IP), ("Xpts", _DP), ("thick", _DP),
                ("temp", _DP), ("mat", C.POINTER(Material)), ("nbc", C.c_int), ("bc_nodes", _IP),
                ("bc_mask", _IP), ("force", _DP), ("sigma", C.c_double), ("max_lanczos", C.c_int),
                ("nev", C.c_int), ("seed", C.c_ulonglong), ("u0", _DP), ("bc_vals", _DP)]


def tridiag_eig(d, e):
    d = _f64(d)
    e = _f64(e)
    n = d.size
    eig = np.zeros(n)
    z = np.zeros(n * n)
    rc = lib().or_tridiag_eig(n, _d(d), _d(e), _d(eig), _d(z))
    if rc != 0:
        raise RuntimeError("QL did not converge")
    return eig, z.reshape(n, n)


def linear_buckling(Xpts, conn, thick, bcs, force, sigma, nev=5, max_lanczos=60, mat=None, temp=None, seed=12345,
                    u0=None):
    """Oracle linear buckling: returns (lambda[nev], phi[nev, 6n], path[6n],
    resid[nev], nsteps).  u0 (optional) is the path; otherwise the static
    solution of K u = f.  bcs = (nodes, masks) or (nodes, masks, values[6
    per node]) for prescribed displacements."""
    mat = mat or material()
    conn = np.ascontiguousarray(conn, np.int32)
    Xpts = _f64(Xpts)
    thick = _f64(thick)
    bn = np.ascontiguousarray(bcs[0], np.int32)
    bm = np.ascontiguousarray(bcs[1], np.int32)
    force = _f64(force)
    tp = _f64(temp) if temp is not None else None
    nn = Xpts.size // 3
    u0 = _f64(u0) if u0 is not None else None
    bv = _f64(bcs[2]) if len(bcs) > 2 and bcs[2] is not None else None
    p = BucklingProblem(nn, conn.size // 4, _i(conn), _d(Xpts), _d(thick), _d(tp) if tp is not None else None,
                        C.pointer(mat), bn.size, _i(bn), _i(bm), _d(force), sigma, max_lanczos, nev, seed,
                        _d(u0) if u0 is not None else None, _d(bv) if bv is not None else None)
    lam = np.zeros(nev)
    phi = np.zeros((nev, 6 * nn))
    path = np.zeros(6 * nn)
    resid = np.zeros(nev)
    ns = lib().or_linear_buckling(C.byref(p), _d(lam), _d(phi), _d(path), _d(resid))
    if ns < 0:
        raise RuntimeError(f"oracle buckling failed ({ns})")
    return lam, phi, path, resid, ns


def rcm(rowp, cols):
    rowp = np.ascontiguousarray(rowp, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    perm = np.empty(rowp.size - 1, np.int32)
    lib().or_rcm(rowp.size - 1, _i(rowp), _i(cols), _i(perm))
    return perm


def band_solve(rowp, cols, vals, b):
    """Direct solve A x = b with the oracle's RCM-ordered band LU."""
    rowp = np.ascontiguousarray(rowp, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    vals = _f64(vals)
    perm = rcm(rowp, cols)
    f = lib().or_band_factor
    f.restype = C.c_void_p
    h = f(rowp.size - 1, _i(rowp), _i(cols), _d(vals), _i(perm))
    x = np.zeros(6 * (rowp.size - 1))
    lib().or_band_solve(C.c_void_p(h), _d(_f64(b)), _d(x))
    lib().or_band_free(C.c_void_p(h))
    return x
