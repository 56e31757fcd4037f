"""Synthetic shell problems (host side): the axially compressed cylinder and
its seeded thickness / imperfection fields.

Node numbering follows the reference's TACSCreator rule (nodes numbered in
order of first appearance in the element connectivity,
src/TACSCreator.cpp:1141-1202) so that patterns and orderings match what the
reference would build from the same connectivity.
"""
import numpy as np

E_ALU, NU_ALU, RHO_ALU = 70e9, 0.3, 2700.0


def first_appearance(conn, nnodes):
    """new_of_old[i]: TACSCreator's first-appearance node numbering."""
    new_of_old = np.full(nnodes, -1, dtype=np.int64)
    nxt = 0
    for n in np.asarray(conn).ravel():
        if new_of_old[n] < 0:
            new_of_old[n] = nxt
            nxt += 1
    return new_of_old


class Cylinder:
    """Thin cylinder of radius R and length L along x, na axial x nc
    circumferential 4-node quads, periodic in theta.

    Elements are ordered with the axial index fastest; element e = j*na + i
    has nodes [(i,j), (i+1,j), (i,j+1), (i+1,j+1)] (tensor order of the
    reference's quad basis).  BCs: x = 0 ring clamped (mask 0x3F), x = L ring
    clamped except the axial displacement (mask 0x3E).  Load: axial
    compressive edge load `line_load` (N/m) lumped on the x = L nodes.
    """

    def __init__(self, na, nc, R=1.0, L=2.0, imperfection=None):
        self.na, self.nc, self.R, self.L = na, nc, R, L
        grid = np.arange((na + 1) * nc).reshape(nc, na + 1).T  # grid[i, j]
        conn = np.empty((na * nc, 4), dtype=np.int64)
        e = 0
        for j in range(nc):
            for i in range(na):
                jp = (j + 1) % nc
                conn[e] = [grid[i, j], grid[i + 1, j], grid[i, jp], grid[i + 1, jp]]
                e += 1
        nnodes = (na + 1) * nc
        perm = first_appearance(conn, nnodes)
        self.conn = perm[conn].astype(np.int32)
        ii, jj = np.meshgrid(np.arange(na + 1), np.arange(nc), indexing="ij")
        x = ii * (L / na)
        th = 2.0 * np.pi * jj / nc
        r = np.full_like(x, R, dtype=float)
        if imperfection is not None:
            amp, m, n = imperfection
            r = r + amp * np.sin(m * np.pi * x / L) * np.cos(n * th)
        X = np.stack([x, r * np.cos(th), r * np.sin(th)], axis=-1)  # [i, j, 3]
        Xpts = np.empty((nnodes, 3))
        Xpts[perm[grid.ravel()]] = X.reshape(-1, 3)
        self.Xpts = Xpts
        self.nnodes, self.nelems = nnodes, na * nc
        self.axial_index = np.empty(nnodes, dtype=np.int64)
        self.axial_index[perm[grid.ravel()]] = ii.ravel()

    def bcs(self):
        n0 = np.where(self.axial_index == 0)[0]
        n1 = np.where(self.axial_index == self.na)[0]
        nodes = np.concatenate([n0, n1]).astype(np.int32)
        mask = np.concatenate([np.full(n0.size, 0x3F), np.full(n1.size, 0x3E)]).astype(np.int32)
        return nodes, mask

    def ring_nodes(self, i):
        """Node ids of axial station i (a seed for the level ordering)."""
        return np.where(self.axial_index == i)[0].astype(np.int32)

    def classical_load(self, t=0.002, E=E_ALU, nu=NU_ALU):
        """Classical axial buckling line load E t^2 / (R sqrt(3(1-nu^2)))
        (N/m), an estimate used to place the shift."""
        return E * t * t / (self.R * np.sqrt(3.0 * (1.0 - nu * nu)))

    def edge_load(self, line_load=1.0):
        f = np.zeros((self.nnodes, 6))
        n1 = np.where(self.axial_index == self.na)[0]
        f[n1, 0] = -line_load * 2.0 * np.pi * self.R / self.nc
        return f.ravel()


def seeded_thickness(nelems, seed, lo=0.0015, hi=0.0025, ndesigns=1):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=(ndesigns, nelems))


def seeded_imperfection(seed, tbar=0.002):
    rng = np.random.default_rng(seed)
    m, n = rng.integers(1, 9, size=2)
    return (0.1 * tbar, int(m), int(n))
