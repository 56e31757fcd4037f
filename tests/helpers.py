"""Shared test helpers: parity metrics and problem builders."""
import numpy as np


def row_floor_ratio(test, ref, floor=1e-3):
    """max_ij |test_ij - ref_ij| / max(|ref_ij|, floor * max_j |ref_ij|),
    rows being scalar rows of the (possibly BSR-stored) matrix.

    This is the entry-wise relative metric with a row-scale floor used for
    the assembled K and G parity (a relative error with a floor of 1e-3 of
    the largest entry of the row, so structurally tiny entries do not divide
    by ~0)."""
    t = np.asarray(test, dtype=float)
    r = np.asarray(ref, dtype=float)
    if r.ndim == 3:  # (nnzb, 6, 6) BSR values -> per-block rows; caller passes rows
        raise ValueError("pass dense matrices or use bsr_row_floor_ratio")
    rowmax = np.abs(r).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(r), floor * rowmax)
    den[den == 0] = 1.0
    return float((np.abs(t - r) / den).max())


def bsr_row_floor_ratio(rowp, test_vals, ref_vals, floor=1e-3):
    t = np.asarray(test_vals).reshape(-1, 6, 6)
    r = np.asarray(ref_vals).reshape(-1, 6, 6)
    worst = 0.0
    n = len(rowp) - 1
    # scalar-row maxima of the reference
    rowmax = np.zeros((n, 6))
    for i in range(n):
        blk = np.abs(r[rowp[i]:rowp[i + 1]])
        if blk.size:
            rowmax[i] = blk.max(axis=(0, 2))
    rm = np.repeat(rowmax, np.diff(rowp), axis=0)[:, :, None]  # (nnzb, 6, 1)
    den = np.maximum(np.abs(r), floor * rm)
    den[den == 0] = 1.0
    worst = float((np.abs(t - r) / den).max())
    return worst


def random_mesh(nx, ny, seed, jitter=0.08, curve=0.3):
    """A small structured mesh of distorted, curved quads (not periodic)."""
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.linspace(0, 1, nx + 1), np.linspace(0, 0.7, ny + 1), indexing="ij")
    X = np.stack([xs, ys, curve * (xs - 0.5) ** 2 + 0.2 * xs * ys], axis=-1)
    X[1:-1, 1:-1] += jitter / max(nx, ny) * rng.standard_normal((nx - 1, ny - 1, 3))
    nid = np.arange((nx + 1) * (ny + 1)).reshape(nx + 1, ny + 1)
    conn = []
    for j in range(ny):
        for i in range(nx):
            conn.append([nid[i, j], nid[i + 1, j], nid[i, j + 1], nid[i + 1, j + 1]])
    return np.array(conn, dtype=np.int32), X.reshape(-1, 3)
