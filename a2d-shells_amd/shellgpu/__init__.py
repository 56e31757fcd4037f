"""Python binding of libshellgpu (the MI355X shell assembly / buckling library).

Thin ctypes layer over the C ABI declared in include/shellgpu.h.  The library
is loaded from ``a2d-shells_amd/lib/libshellgpu.so`` (built in-tree by
``make -C a2d-shells_amd``); there is no CPU fallback: a missing library or a
failing call raises :class:`ShellGPUError`.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libshellgpu.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_ROOT), "include", "shellgpu.h")

STIFFNESS, GEOMETRIC, GEOMETRIC_FD, TANGENT_NONLINEAR, MASS = 0, 1, 2, 3, 4
LINEAR, NONLINEAR = 0, 1


class ShellGPUError(RuntimeError):
    pass


_lib = None
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_VP = C.c_void_p


def load_library(path=LIB_PATH):
    """Load libshellgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ShellGPUError(
            f"{path} not found: build it with `make -C a2d-shells_amd` "
            "(or __graft_entry__.build())")
    lib = C.CDLL(path)
    sig = {
        "shellgpu_create": (C.c_int, [C.POINTER(_VP), C.c_int]),
        "shellgpu_destroy": (None, [_VP]),
        "shellgpu_last_error": (C.c_char_p, [_VP]),
        "shellgpu_synchronize": (C.c_int, [_VP]),
        "shellgpu_get_stream": (_VP, [_VP]),
        "shellgpu_malloc": (C.c_int, [_VP, C.POINTER(_VP), C.c_size_t]),
        "shellgpu_free": (C.c_int, [_VP, _VP]),
        "shellgpu_memcpy_h2d": (C.c_int, [_VP, _VP, _VP, C.c_size_t]),
        "shellgpu_memcpy_d2h": (C.c_int, [_VP, _VP, _VP, C.c_size_t]),
        "shellgpu_memcpy_d2d": (C.c_int, [_VP, _VP, _VP, C.c_size_t]),
        "shellgpu_memset": (C.c_int, [_VP, _VP, C.c_int, C.c_size_t]),
        "shellgpu_set_mesh": (C.c_int, [_VP, C.c_int, C.c_int, _IP, _DP]),
        "shellgpu_set_nodes": (C.c_int, [_VP, _VP]),
        "shellgpu_get_nodes": (C.c_int, [_VP, _VP]),
        "shellgpu_set_material": (C.c_int, [_VP] + [C.c_double] * 5),
        "shellgpu_set_bcs": (C.c_int, [_VP, C.c_int, _IP, _IP, _DP]),
        "shellgpu_set_designs": (C.c_int, [_VP, C.c_int, _DP, _DP]),
        "shellgpu_get_sizes": (C.c_int, [_VP] + [_IP] * 5),
        "shellgpu_get_pattern": (C.c_int, [_VP, _IP, _IP]),
        "shellgpu_get_colors": (C.c_int, [_VP, _IP]),
        "shellgpu_mat_create": (C.c_int, [_VP, C.c_int, C.POINTER(_VP)]),
        "shellgpu_mat_destroy": (None, [_VP]),
        "shellgpu_mat_values": (_VP, [_VP]),
        "shellgpu_mat_get_values": (C.c_int, [_VP, _DP]),
        "shellgpu_mat_set_values": (C.c_int, [_VP, _DP]),
        "shellgpu_mat_zero": (C.c_int, [_VP]),
        "shellgpu_mat_copy": (C.c_int, [_VP, _VP]),
        "shellgpu_mat_axpy": (C.c_int, [_VP, C.c_double, _VP]),
        "shellgpu_mat_apply_bcs": (C.c_int, [_VP, C.c_double]),
        "shellgpu_mat_mult": (C.c_int, [_VP, _VP, _VP]),
        "shellgpu_mat_mult_transpose": (C.c_int, [_VP, _VP, _VP]),
        "shellgpu_mat_scale": (C.c_int, [_VP, C.c_double]),
        "shellgpu_mat_add_diag": (C.c_int, [_VP, C.c_double]),
        "shellgpu_mat_add_blocks": (C.c_int, [_VP, C.c_int, C.c_int, _IP, _IP, _DP]),
        "shellgpu_assemble_mat_combo": (C.c_int, [_VP, C.c_int, _IP, _DP, _VP, _VP]),
        "shellgpu_add_jacobian_vec_product": (C.c_int, [_VP, C.c_int] + [C.c_double] * 4 + [_VP] * 5 + [C.c_int]),
        "shellgpu_assemble_mattype": (C.c_int, [_VP, C.c_int, _VP, _VP]),
        "shellgpu_assemble_res": (C.c_int, [_VP, C.c_int, _VP, _VP]),
        "shellgpu_assemble_jacobian": (C.c_int, [_VP, C.c_int] + [C.c_double] * 3 + [_VP] * 5),
        "shellgpu_add_adj_res_product": (C.c_int, [_VP, C.c_int, C.c_double, _VP, _VP, _VP]),
        "shellgpu_add_mat_dv_sens_inner_product": (C.c_int, [_VP, C.c_int, C.c_double, _VP, _VP, _VP, _VP]),
        "shellgpu_eval_mat_sv_sens_inner_product": (C.c_int, [_VP, _VP, _VP, _VP]),
        "shellgpu_add_mat_xpt_sens_inner_product": (C.c_int, [_VP, C.c_int, C.c_double, _VP, _VP, _VP, _VP]),
        "shellgpu_add_adj_res_xpt_product": (C.c_int, [_VP, C.c_int, C.c_double, _VP, _VP, _VP]),
        "shellgpu_vec_zero_bcs": (C.c_int, [_VP, _VP]),
        "shellgpu_vec_set_bcs": (C.c_int, [_VP, _VP]),
        "shellgpu_set_ordering_seed": (C.c_int, [_VP, C.c_int, _IP]),
        "shellgpu_factor_create": (C.c_int, [_VP, C.POINTER(_VP)]),
        "shellgpu_factor_destroy": (None, [_VP]),
        "shellgpu_factor_info": (C.c_int, [_VP, _IP, _IP, _DP]),
        "shellgpu_factor_compute": (C.c_int, [_VP, _VP]),
        "shellgpu_factor_apply": (C.c_int, [_VP, _VP, _VP]),
        "shellgpu_gmres": (C.c_int, [_VP, _VP, _VP, _VP, _VP, C.c_int, C.c_int, C.c_double, C.c_double, _IP, _DP]),
        "shellgpu_gmres_noalloc": (C.c_int, [_VP, _VP, _VP, _VP, _VP, C.c_int, C.c_int, C.c_double, C.c_double, _IP,
                                             _DP]),
        "shellgpu_linear_buckling": (C.c_int, [_VP, C.c_double, _VP, C.c_int, C.c_int, C.c_double, _DP, _VP, _VP,
                                               _IP, _DP, _DP]),
        "shellgpu_buckling_create": (C.c_int, [_VP, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                               C.POINTER(_VP)]),
        "shellgpu_buckling_destroy": (None, [_VP]),
        "shellgpu_buckling_solve": (C.c_int, [_VP, _VP, _VP]),
        "shellgpu_buckling_extract": (C.c_int, [_VP, C.c_int, _DP, _VP, _DP]),
        "shellgpu_buckling_info": (C.c_int, [_VP, _IP, _DP, _VP]),
        "shellgpu_buckling_add_eigen_sens": (C.c_int, [_VP, _DP, _VP, _VP]),
        "shellgpu_buckling_set_stiffness_factor": (C.c_int, [_VP, _VP]),
        "shellgpu_newton_solve_factor": (C.c_int, [_VP, _VP, _VP, C.c_int, C.c_int, C.c_double, _VP, _IP, _DP]),
        "shellgpu_element_output": (C.c_int, [_VP, C.c_int, _VP, C.c_double, _VP]),
        "shellgpu_frequency_create": (C.c_int, [_VP, C.c_double, C.c_int, C.c_int, C.c_double, C.POINTER(_VP)]),
        "shellgpu_frequency_destroy": (None, [_VP]),
        "shellgpu_frequency_solve": (C.c_int, [_VP]),
        "shellgpu_frequency_extract": (C.c_int, [_VP, C.c_int, _DP, _VP, _DP]),
        "shellgpu_frequency_info": (C.c_int, [_VP, _IP, _DP]),
        "shellgpu_frequency_add_eigen_sens": (C.c_int, [_VP, _DP, _VP, _VP]),
        "shellgpu_ks_aggregate": (C.c_int, [C.c_int, C.c_double, _DP, _DP, _DP]),
        "shellgpu_newton_solve": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_double, _VP, _IP, _DP]),
        "shellgpu_device_info": (C.c_int, [_VP, C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_int, _IP,
                                           C.POINTER(C.c_size_t)]),
        "shellgpu_tune": (C.c_int, [_VP, C.c_int]),
        "shellgpu_get_schedule": (C.c_int, [_VP, C.c_char_p, C.c_int]),
        "shellgpu_set_elem_schedule": (C.c_int, [_VP, C.c_int, C.c_int]),
        "shellgpu_set_timing": (C.c_int, [_VP, C.c_int]),
        "shellgpu_get_timing": (C.c_int, [_VP, _DP, _IP, _DP, _IP]),
        "shellgpu_timer_start": (C.c_int, [_VP]),
        "shellgpu_timer_stop": (C.c_int, [_VP, _DP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


class DeviceArray:
    """A device buffer owned by a Context (float64 or int32)."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = _VP()
        ctx._check(ctx.lib.shellgpu_malloc(ctx.h, C.byref(p), max(1, self.n) * self.dtype.itemsize))
        self.ptr = p.value

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype).ravel()
        if host.size != self.n:
            raise ShellGPUError(f"size mismatch {host.size} != {self.n}")
        self.ctx._check(self.ctx.lib.shellgpu_memcpy_h2d(self.ctx.h, self.ptr, host.ctypes.data, self.nbytes))
        return self

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        self.ctx._check(self.ctx.lib.shellgpu_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self):
        self.ctx._check(self.ctx.lib.shellgpu_memset(self.ctx.h, self.ptr, 0, self.nbytes))
        return self

    def free(self):
        if self.ptr:
            self.ctx.lib.shellgpu_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


class Matrix:
    """A batch of BSR matrices sharing the context's pattern (TACSMat)."""

    def __init__(self, ctx, nbatch=None):
        self.ctx = ctx
        self.nbatch = nbatch or ctx.nbatch
        h = _VP()
        ctx._check(ctx.lib.shellgpu_mat_create(ctx.h, self.nbatch, C.byref(h)))
        self.h = h.value

    @property
    def device_values(self):
        return self.ctx.lib.shellgpu_mat_values(self.h)

    def values(self):
        out = np.empty(self.nbatch * self.ctx.nnzb * 36)
        self.ctx._check(self.ctx.lib.shellgpu_mat_get_values(self.h, _dp(out)))
        return out.reshape(self.nbatch, self.ctx.nnzb, 6, 6)

    def design_values(self, d):
        """Values of design d only, (nnzb, 6, 6)."""
        if not 0 <= d < self.nbatch:
            raise ShellGPUError(f"design {d} out of range")
        n = self.ctx.nnzb * 36
        out = np.empty(n)
        self.ctx.synchronize()
        self.ctx._check(self.ctx.lib.shellgpu_memcpy_d2h(self.ctx.h, out.ctypes.data, self.device_values + 8 * n * d,
                                                         8 * n))
        return out.reshape(self.ctx.nnzb, 6, 6)

    def set_values(self, vals):
        vals = np.ascontiguousarray(vals, dtype=np.float64).ravel()
        self.ctx._check(self.ctx.lib.shellgpu_mat_set_values(self.h, _dp(vals)))

    def zero(self):
        self.ctx._check(self.ctx.lib.shellgpu_mat_zero(self.h))

    def copy_from(self, other):
        self.ctx._check(self.ctx.lib.shellgpu_mat_copy(self.h, other.h))

    def axpy(self, alpha, other):
        self.ctx._check(self.ctx.lib.shellgpu_mat_axpy(self.h, float(alpha), other.h))

    def apply_bcs(self, diag=1.0):
        self.ctx._check(self.ctx.lib.shellgpu_mat_apply_bcs(self.h, float(diag)))

    def mult(self, x, y):
        self.ctx._check(self.ctx.lib.shellgpu_mat_mult(self.h, x.ptr, y.ptr))

    def mult_transpose(self, x, y):
        self.ctx._check(self.ctx.lib.shellgpu_mat_mult_transpose(self.h, x.ptr, y.ptr))

    def scale(self, alpha):
        self.ctx._check(self.ctx.lib.shellgpu_mat_scale(self.h, float(alpha)))

    def add_diag(self, alpha):
        self.ctx._check(self.ctx.lib.shellgpu_mat_add_diag(self.h, float(alpha)))

    def add_blocks(self, rows, cols, vals, design=0):
        """TACSMat::addValues as blocks: vals[k] (6x6, row-major) added to
        block (rows[k], cols[k]) of design `design`."""
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64).reshape(-1, 36)
        self.ctx._check(self.ctx.lib.shellgpu_mat_add_blocks(self.h, int(design), rows.size, _ip(rows), _ip(cols),
                                                             _dp(vals)))

    def destroy(self):
        if self.h and self.ctx.h:
            self.ctx.lib.shellgpu_mat_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Factor:
    """Level-set block LU of a batch of matrices (TACSPc)."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = _VP()
        ctx._check(ctx.lib.shellgpu_factor_create(ctx.h, C.byref(h)))
        self.h = h.value

    def info(self):
        nl, mx, gib = C.c_int(), C.c_int(), C.c_double()
        self.ctx._check(self.ctx.lib.shellgpu_factor_info(self.h, C.byref(nl), C.byref(mx), C.byref(gib)))
        return nl.value, mx.value, gib.value

    def compute(self, A):
        self.ctx._check(self.ctx.lib.shellgpu_factor_compute(self.h, A.h))

    def apply(self, b, x):
        self.ctx._check(self.ctx.lib.shellgpu_factor_apply(self.h, b.ptr, x.ptr))

    def destroy(self):
        if self.h and self.ctx.h:
            self.ctx.lib.shellgpu_factor_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Frequency:
    """Natural-frequency analysis of every design of a context
    (TACSFrequencyAnalysis, Lanczos path): K x = lambda M x, lambda = omega^2."""

    def __init__(self, ctx, sigma=0.0, nev=5, max_lanczos=100, tol=1e-12):
        self.ctx = ctx
        self.nev = nev
        h = _VP()
        ctx._check(ctx.lib.shellgpu_frequency_create(ctx.h, float(sigma), nev, max_lanczos, tol, C.byref(h)))
        self.h = h.value

    def solve(self):
        self.ctx._check(self.ctx.lib.shellgpu_frequency_solve(self.h))

    def eigenvalues(self):
        """[nbatch, nev] eigenvalues omega^2 and relative Ritz residuals."""
        nb = self.ctx.nbatch
        lam, err = np.zeros((self.nev, nb)), np.zeros((self.nev, nb))
        for j in range(self.nev):
            self.ctx._check(self.ctx.lib.shellgpu_frequency_extract(self.h, j, _dp(lam[j]), None, _dp(err[j])))
        return lam.T.copy(), err.T.copy()

    def eigenvector(self, j):
        """M-normalised eigenvector j of every design, [nbatch, 6n]."""
        nb, n = self.ctx.nbatch, 6 * self.ctx.nnodes
        phi = DeviceArray(self.ctx, nb * n)
        lam = np.zeros(nb)
        self.ctx._check(self.ctx.lib.shellgpu_frequency_extract(self.h, j, _dp(lam), phi.ptr, None))
        return phi.download().reshape(nb, n)

    def info(self):
        steps, sig = C.c_int(), np.zeros(self.ctx.nbatch)
        self.ctx._check(self.ctx.lib.shellgpu_frequency_info(self.h, C.byref(steps), _dp(sig)))
        return {"steps": steps.value, "sigma": sig}

    def eigen_sens(self, coef, xpts=False):
        """sum_j coef[d, j] d(lambda_j)/dt per design, [nbatch, nelems]; with
        xpts also the coordinate derivative, [nbatch, nnodes, 3]."""
        ctx = self.ctx
        nb = ctx.nbatch
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(coef, dtype=np.float64), (nb, self.nev)))
        dt = DeviceArray(ctx, nb * ctx.nelems)
        dt.zero()
        dX = None
        if xpts:
            dX = DeviceArray(ctx, nb * 3 * ctx.nnodes)
            dX.zero()
        ctx._check(ctx.lib.shellgpu_frequency_add_eigen_sens(self.h, _dp(c), dt.ptr, dX.ptr if dX else None))
        g = dt.download().reshape(nb, ctx.nelems)
        if xpts:
            return g, dX.download().reshape(nb, ctx.nnodes, 3)
        return g

    def destroy(self):
        if self.h and self.ctx.h:
            self.ctx.lib.shellgpu_frequency_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Buckling:
    """Linear buckling analysis of every design of a context
    (TACSLinearBuckling): solve, extract modes, eigenvalue sensitivities."""

    def __init__(self, ctx, sigma=0.0, nev=5, max_lanczos=100, tol=1e-12, model=None):
        self.ctx = ctx
        self.nev = nev
        h = _VP()
        ctx._check(ctx.lib.shellgpu_buckling_create(ctx.h, LINEAR if model is None else model, float(sigma), nev,
                                                    max_lanczos, tol, C.byref(h)))
        self.h = h.value

    def set_stiffness_factor(self, F):
        """Share the caller's factor of K (zero state) for the shift
        estimate and the path solves; F must stay alive and current."""
        self.ctx._check(self.ctx.lib.shellgpu_buckling_set_stiffness_factor(self.h, F.h))
        self._fk = F

    def solve(self, force=None, u0=None):
        """Static path (or u0), G, shift, Lanczos; force/u0 are host arrays
        broadcast to [nbatch, 6n] or DeviceArrays."""
        ctx = self.ctx
        nb, n = ctx.nbatch, 6 * ctx.nnodes

        def dev(v):
            if v is None or isinstance(v, DeviceArray):
                return v
            return ctx.vector(np.broadcast_to(v, (nb, n)))
        f, u = dev(force), dev(u0)
        ctx._check(ctx.lib.shellgpu_buckling_solve(self.h, f.ptr if f else None, u.ptr if u else None))

    def eigenvalues(self):
        """[nbatch, nev] eigenvalues and relative Ritz residuals."""
        nb = self.ctx.nbatch
        lam, err = np.zeros((self.nev, nb)), np.zeros((self.nev, nb))
        for j in range(self.nev):
            self.ctx._check(self.ctx.lib.shellgpu_buckling_extract(self.h, j, _dp(lam[j]), None, _dp(err[j])))
        return lam.T.copy(), err.T.copy()

    def eigenvector(self, j):
        """K-normalised eigenvector j of every design, [nbatch, 6n]."""
        nb, n = self.ctx.nbatch, 6 * self.ctx.nnodes
        phi = DeviceArray(self.ctx, nb * n)
        lam = np.zeros(nb)
        self.ctx._check(self.ctx.lib.shellgpu_buckling_extract(self.h, j, _dp(lam), phi.ptr, None))
        return phi.download().reshape(nb, n)

    def info(self):
        nb, n = self.ctx.nbatch, 6 * self.ctx.nnodes
        steps, sig = C.c_int(), np.zeros(nb)
        path = DeviceArray(self.ctx, nb * n)
        self.ctx._check(self.ctx.lib.shellgpu_buckling_info(self.h, C.byref(steps), _dp(sig), path.ptr))
        return {"steps": steps.value, "sigma": sig, "path": path.download().reshape(nb, n)}

    def eigen_sens(self, coef, thickness=True, coords=False):
        """sum_j coef[d, j] d(lambda_j)/dx per design: (dfdt [nbatch,
        nelems] or None, dfdX [nbatch, 3*nnodes] or None)."""
        ctx = self.ctx
        nb = ctx.nbatch
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(coef, dtype=np.float64), (nb, self.nev)))
        dt = DeviceArray(ctx, nb * ctx.nelems) if thickness else None
        dX = DeviceArray(ctx, nb * 3 * ctx.nnodes) if coords else None
        if dt is not None:
            dt.zero()
        if dX is not None:
            dX.zero()
        ctx._check(ctx.lib.shellgpu_buckling_add_eigen_sens(self.h, _dp(c), dt.ptr if dt else None,
                                                           dX.ptr if dX else None))
        return (dt.download().reshape(nb, ctx.nelems) if dt else None,
                dX.download().reshape(nb, 3 * ctx.nnodes) if dX else None)

    def destroy(self):
        if self.h and self.ctx.h:
            self.ctx.lib.shellgpu_buckling_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def ks_aggregate(f, rho):
    """KS aggregate (max-like) of f and its gradient d ks / d f."""
    lib = load_library()
    f = np.ascontiguousarray(f, dtype=np.float64)
    ks, dks = C.c_double(), np.zeros_like(f)
    if lib.shellgpu_ks_aggregate(len(f), float(rho), _dp(f), C.byref(ks), _dp(dks)) != 0:
        raise ShellGPUError("ks_aggregate: invalid arguments")
    return ks.value, dks


class Context:
    """One GPU context: mesh, material, BCs, designs and a HIP stream."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = _VP()
        rc = self.lib.shellgpu_create(C.byref(h), int(device))
        if rc != 0:
            raise ShellGPUError(f"shellgpu_create failed ({rc})")
        self.h = h.value
        self.nnodes = self.nelems = self.nnzb = self.ncolors = self.nbatch = 0

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.shellgpu_last_error(self.h)
            raise ShellGPUError(f"shellgpu error {rc}: {msg.decode() if msg else ''}")

    def _refresh(self):
        v = [C.c_int() for _ in range(5)]
        self._check(self.lib.shellgpu_get_sizes(self.h, *[C.byref(x) for x in v]))
        self.nnodes, self.nelems, self.nnzb, self.ncolors, self.nbatch = [x.value for x in v]

    @property
    def stream(self):
        return self.lib.shellgpu_get_stream(self.h)

    def synchronize(self):
        self._check(self.lib.shellgpu_synchronize(self.h))

    def tune(self, coloring=True):
        """Autotune the element-kernel schedules (and the colouring) on this
        context's mesh and designs; returns the schedule report."""
        self._check(self.lib.shellgpu_tune(self.h, 1 if coloring else 0))
        self._refresh()
        return self.schedule()

    def schedule(self):
        import json
        buf = C.create_string_buffer(8192)
        self._check(self.lib.shellgpu_get_schedule(self.h, buf, 8192))
        return json.loads(buf.value.decode())

    def set_elem_schedule(self, kernel_class, variant):
        self._check(self.lib.shellgpu_set_elem_schedule(self.h, int(kernel_class), int(variant)))

    def device_info(self):
        """The context's GPU: name, architecture, PCI bus id, compute units,
        HBM bytes."""
        name, arch, pci = C.create_string_buffer(256), C.create_string_buffer(64), C.create_string_buffer(64)
        cu, mem = C.c_int(), C.c_size_t()
        self._check(self.lib.shellgpu_device_info(self.h, name, 256, arch, 64, pci, 64, C.byref(cu), C.byref(mem)))
        return {"name": name.value.decode(), "arch": arch.value.decode(), "pci_bus_id": pci.value.decode(),
                "compute_units": cu.value, "hbm_bytes": mem.value}

    def set_mesh(self, conn, Xpts):
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        Xpts = np.ascontiguousarray(Xpts, dtype=np.float64)
        self._check(self.lib.shellgpu_set_mesh(self.h, Xpts.size // 3, conn.size // 4, _ip(conn), _dp(Xpts)))
        self._refresh()

    def set_nodes(self, Xpts):
        """Move the mesh: new nodal coordinates [nnodes, 3], same connectivity."""
        Xpts = np.ascontiguousarray(Xpts, dtype=np.float64)
        if Xpts.size != 3 * self.nnodes:
            raise ShellGPUError("Xpts must hold 3 coordinates per node")
        d = DeviceArray(self, Xpts.size)
        d.upload(Xpts.ravel())
        self._check(self.lib.shellgpu_set_nodes(self.h, d.ptr))
        self.synchronize()

    def get_nodes(self):
        d = DeviceArray(self, 3 * self.nnodes)
        self._check(self.lib.shellgpu_get_nodes(self.h, d.ptr))
        return d.download().reshape(-1, 3)

    def set_material(self, E, nu, rho=2700.0, cte=0.0, k_drill=10.0):
        self._check(self.lib.shellgpu_set_material(self.h, E, nu, rho, cte, k_drill))

    def set_bcs(self, nodes, mask, vals=None):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        vp = None
        if vals is not None:
            vals = np.ascontiguousarray(vals, dtype=np.float64)
            vp = _dp(vals)
        self._check(self.lib.shellgpu_set_bcs(self.h, nodes.size, _ip(nodes), _ip(mask), vp))

    def set_designs(self, thick, temp=None):
        thick = np.ascontiguousarray(np.atleast_2d(thick), dtype=np.float64)
        nb = thick.shape[0]
        tp = None
        if temp is not None:
            temp = np.ascontiguousarray(np.broadcast_to(temp, thick.shape), dtype=np.float64)
            tp = _dp(temp)
        self._check(self.lib.shellgpu_set_designs(self.h, nb, _dp(thick), tp))
        self._refresh()

    def pattern(self):
        rowp = np.empty(self.nnodes + 1, np.int32)
        cols = np.empty(self.nnzb, np.int32)
        self._check(self.lib.shellgpu_get_pattern(self.h, _ip(rowp), _ip(cols)))
        return rowp, cols

    def colors(self):
        c = np.empty(self.nelems, np.int32)
        self._check(self.lib.shellgpu_get_colors(self.h, _ip(c)))
        return c

    def vector(self, host=None, nbatch=None):
        n = 6 * self.nnodes * (nbatch or max(self.nbatch, 1))
        d = DeviceArray(self, n)
        if host is None:
            d.zero()
        else:
            d.upload(host)
        return d

    def matrix(self, nbatch=None):
        return Matrix(self, nbatch)

    def assemble_mattype(self, mattype, A, u=None):
        self._check(self.lib.shellgpu_assemble_mattype(self.h, int(mattype), u.ptr if u else None, A.h))

    def assemble_res(self, R, u=None, model=LINEAR):
        self._check(self.lib.shellgpu_assemble_res(self.h, int(model), u.ptr if u else None, R.ptr))

    def assemble_jacobian(self, A, R=None, u=None, model=LINEAR, alpha=1.0, beta=0.0, gamma=0.0, dudt=None,
                          d2udt2=None):
        """assembleJacobian: A = alpha dR/du + gamma M, R = R(u) + M d2udt2
        (BCs applied to both)."""
        self._check(self.lib.shellgpu_assemble_jacobian(
            self.h, int(model), float(alpha), float(beta), float(gamma), u.ptr if u else None,
            dudt.ptr if dudt else None, d2udt2.ptr if d2udt2 else None, R.ptr if R else None, A.h))

    def assemble_mat_combo(self, A, mattypes, scale, u=None):
        t = np.ascontiguousarray(mattypes, dtype=np.int32)
        c = np.ascontiguousarray(scale, dtype=np.float64)
        self._check(self.lib.shellgpu_assemble_mat_combo(self.h, t.size, _ip(t), _dp(c), u.ptr if u else None, A.h))

    def add_jacobian_vec_product(self, y, x, scale=1.0, alpha=1.0, beta=0.0, gamma=0.0, u=None, model=LINEAR,
                                 transpose=False):
        """y += scale (alpha dR/du + gamma M) x (or its transpose), then
        y[bc] = 0 (addJacobianVecProduct)."""
        self._check(self.lib.shellgpu_add_jacobian_vec_product(
            self.h, int(model), float(scale), float(alpha), float(beta), float(gamma), u.ptr if u else None,
            None, None, x.ptr, y.ptr, int(bool(transpose))))

    def set_ordering_seed(self, nodes):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        self._check(self.lib.shellgpu_set_ordering_seed(self.h, nodes.size, _ip(nodes)))

    def factor(self):
        return Factor(self)

    def gmres(self, A, b, x, F=None, msub=10, nrestart=15, rtol=1e-12, atol=1e-30):
        its, rn = C.c_int(), C.c_double()
        self._check(self.lib.shellgpu_gmres(self.h, A.h, F.h if F else None, b.ptr, x.ptr, msub, nrestart, rtol,
                                            atol, C.byref(its), C.byref(rn)))
        return its.value, rn.value

    def gmres_noalloc(self, A, b, x, F=None, msub=10, nrestart=15, rtol=1e-12, atol=1e-30):
        """gmres() on the library's fixed device scratch array (no device
        allocation; the workspace must fit in 64 MiB)"""
        its, rn = C.c_int(), C.c_double()
        self._check(self.lib.shellgpu_gmres_noalloc(self.h, A.h, F.h if F else None, b.ptr, x.ptr, msub, nrestart,
                                                    rtol, atol, C.byref(its), C.byref(rn)))
        return its.value, rn.value

    def linear_buckling(self, sigma, force, nev=5, max_lanczos=100, tol=1e-12, want_vectors=False, want_path=False):
        """TACSLinearBuckling::solve for every design: returns lambda
        [nbatch, nev], Ritz residuals, Lanczos steps, and optionally the
        eigenvectors [nbatch, nev, 6n] and the path [nbatch, 6n]."""
        nb, n = self.nbatch, 6 * self.nnodes
        f = force if isinstance(force, DeviceArray) else self.vector(np.broadcast_to(force, (nb, n)))
        lam = np.zeros(nb * nev)
        res = np.zeros(nb * nev)
        sig = np.zeros(nb)
        steps = C.c_int()
        phi = DeviceArray(self, nb * nev * n) if want_vectors else None
        path = DeviceArray(self, nb * n) if want_path else None
        self._check(self.lib.shellgpu_linear_buckling(self.h, float(sigma), f.ptr, nev, max_lanczos, tol, _dp(lam),
                                                      phi.ptr if phi else None, path.ptr if path else None,
                                                      C.byref(steps), _dp(res), _dp(sig)))
        out = {"lambda": lam.reshape(nb, nev), "resid": res.reshape(nb, nev), "steps": steps.value, "sigma": sig}
        if phi is not None:
            out["phi"] = phi.download().reshape(nb, nev, n)
        if path is not None:
            out["path"] = path.download().reshape(nb, n)
        return out

    def design_vector(self):
        """Zeroed device array of per-design element values [nbatch*nelems]."""
        return DeviceArray(self, self.nbatch * self.nelems).zero()

    def add_adj_res_product(self, dfdt, psi, u=None, scale=1.0, model=LINEAR):
        self._check(self.lib.shellgpu_add_adj_res_product(self.h, model, float(scale), psi.ptr,
                                                          u.ptr if u else None, dfdt.ptr))

    def add_mat_dv_sens_inner_product(self, dfdt, mattype, phi, u=None, scale=1.0, psi=None):
        """dfdt += scale psi^T (dA/dt) phi (psi None: psi = phi)"""
        self._check(self.lib.shellgpu_add_mat_dv_sens_inner_product(self.h, mattype, float(scale),
                                                                    psi.ptr if psi else None, phi.ptr,
                                                                    u.ptr if u else None, dfdt.ptr))

    def coord_vector(self):
        """Zeroed device array of per-design nodal coordinates [nbatch*3*nnodes]."""
        return DeviceArray(self, self.nbatch * 3 * self.nnodes).zero()

    def add_mat_xpt_sens_inner_product(self, dfdX, mattype, phi, u=None, scale=1.0, psi=None):
        """dfdX += scale d(psi^T A phi)/dX (psi None: psi = phi)"""
        self._check(self.lib.shellgpu_add_mat_xpt_sens_inner_product(self.h, mattype, float(scale),
                                                                     psi.ptr if psi else None, phi.ptr,
                                                                     u.ptr if u else None, dfdX.ptr))

    def add_adj_res_xpt_product(self, dfdX, psi, u=None, scale=1.0, model=LINEAR):
        self._check(self.lib.shellgpu_add_adj_res_xpt_product(self.h, model, float(scale), psi.ptr,
                                                              u.ptr if u else None, dfdX.ptr))

    def eval_mat_sv_sens_inner_product(self, phi, dfdu, psi=None):
        """dfdu = d(psi^T G(u) phi)/du (psi None: psi = phi)"""
        self._check(self.lib.shellgpu_eval_mat_sv_sens_inner_product(self.h, psi.ptr if psi else None, phi.ptr,
                                                                     dfdu.ptr))

    def newton_solve(self, force, nsteps=5, max_iters=20, rtol=1e-10, u0=None, factor=None):
        """Load-stepped Newton on the nonlinear model: returns (u device
        array [nbatch*6n], Newton iterations, relative residuals [nbatch]).
        With a Factor, the tangent factors go into it and the last stays
        there (from u0 = 0 at zero temperature: the stiffness K)."""
        nb, n = self.nbatch, 6 * self.nnodes
        f = force if isinstance(force, DeviceArray) else self.vector(np.broadcast_to(force, (nb, n)))
        u = self.vector(None if u0 is None else np.broadcast_to(u0, (nb, n)))
        its, rel = C.c_int(), np.zeros(nb)
        if factor is None:
            self._check(self.lib.shellgpu_newton_solve(self.h, f.ptr, nsteps, max_iters, rtol, u.ptr,
                                                       C.byref(its), _dp(rel)))
        else:
            self._check(self.lib.shellgpu_newton_solve_factor(self.h, factor.h, f.ptr, nsteps, max_iters, rtol,
                                                              u.ptr, C.byref(its), _dp(rel)))
        return u, its.value, rel

    def buckling(self, sigma=0.0, nev=5, max_lanczos=100, tol=1e-12, model=None):
        return Buckling(self, sigma, nev, max_lanczos, tol, model)

    def element_output(self, u=None, model=LINEAR, ys=1e11):
        """Strains, stresses, failure and design values at the element
        nodes, [nbatch, nelems, 4, 31] (getOutputData)."""
        nb = self.nbatch
        du = None
        if u is not None:
            du = u if isinstance(u, DeviceArray) else self.vector(np.broadcast_to(u, (nb, 6 * self.nnodes)))
        out = DeviceArray(self, nb * self.nelems * 4 * 31)
        self._check(self.lib.shellgpu_element_output(self.h, model, du.ptr if du else None, float(ys), out.ptr))
        return out.download().reshape(nb, self.nelems, 4, 31)

    def frequency(self, sigma=0.0, nev=5, max_lanczos=100, tol=1e-12):
        return Frequency(self, sigma, nev, max_lanczos, tol)

    def vec_zero_bcs(self, x):
        self._check(self.lib.shellgpu_vec_zero_bcs(self.h, x.ptr))

    def vec_set_bcs(self, x):
        self._check(self.lib.shellgpu_vec_set_bcs(self.h, x.ptr))

    def set_timing(self, enable=True):
        self._check(self.lib.shellgpu_set_timing(self.h, int(bool(enable))))

    def get_timing(self):
        """(element-kernel ms, launches, spmv ms, launches) since enabling."""
        em, sm_ = C.c_double(), C.c_double()
        el, sl = C.c_int(), C.c_int()
        self._check(self.lib.shellgpu_get_timing(self.h, C.byref(em), C.byref(el), C.byref(sm_), C.byref(sl)))
        return em.value, el.value, sm_.value, sl.value

    def timer_start(self):
        self._check(self.lib.shellgpu_timer_start(self.h))

    def timer_stop(self):
        ms = C.c_double()
        self._check(self.lib.shellgpu_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            self.lib.shellgpu_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
