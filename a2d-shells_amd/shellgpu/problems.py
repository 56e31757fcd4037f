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


class BoxBeam:
    """Stiffened cantilever box beam: span along x, chord along y, height
    along z; top and bottom skins (ns x nc quads), front and rear spars
    (ns x nh), and nribs ribs (nc x nh) at equal spacing ending with one at
    the tip.  Pieces share the nodes of their junction lines, so the model is
    one connected shell with T- and L-junctions.  At the default sizes (400 x
    80 skins, 400 x 16 spars, 40 ribs of 80 x 16, span 10, chord 2, height
    0.4) it has 128,000 quads and 124,392 nodes.

    Panels: each skin and spar is cut by the ribs into nribs bays, each rib
    is one panel (panel_of[e] for element e; 5 nribs panels).  BCs: every
    node of the root station x = 0 clamped (mask 0x3F).  Load cases
    (load_case): a lifting tip force on the tip rib plus a uniform upward
    pressure on the bottom skin, mixed in a proportion set per case.
    Nodes are numbered in order of first appearance in the connectivity.
    """

    def __init__(self, ns=400, nc=80, nh=16, nribs=40, span=10.0, chord=2.0, height=0.4):
        if ns % nribs:
            raise ValueError("the span elements must split evenly into rib bays")
        self.ns, self.nc, self.nh, self.nribs = ns, nc, nh, nribs
        self.span, self.chord, self.height = span, chord, height
        ids = {}
        coords = []

        def node(i, j, k):
            key = (i, j, k)
            n = ids.get(key)
            if n is None:
                n = ids[key] = len(coords)
                coords.append((i * span / ns, j * chord / nc, k * height / nh))
            return n

        conn, panel = [], []
        bay = ns // nribs
        # top skin (outward +z): xi along x, eta along y
        for j in range(nc):
            for i in range(ns):
                conn.append([node(i, j, nh), node(i + 1, j, nh), node(i, j + 1, nh), node(i + 1, j + 1, nh)])
                panel.append(i // bay)
        # bottom skin (outward -z): xi along y, eta along x
        for j in range(nc):
            for i in range(ns):
                conn.append([node(i, j, 0), node(i, j + 1, 0), node(i + 1, j, 0), node(i + 1, j + 1, 0)])
                panel.append(nribs + i // bay)
        # front spar y = 0 (outward -y): xi along x, eta along z
        for k in range(nh):
            for i in range(ns):
                conn.append([node(i, 0, k), node(i + 1, 0, k), node(i, 0, k + 1), node(i + 1, 0, k + 1)])
                panel.append(2 * nribs + i // bay)
        # rear spar y = chord (outward +y): xi along z, eta along x
        for k in range(nh):
            for i in range(ns):
                conn.append([node(i, nc, k), node(i, nc, k + 1), node(i + 1, nc, k), node(i + 1, nc, k + 1)])
                panel.append(3 * nribs + i // bay)
        # ribs at x = r * bay (r = 1..nribs), normal +x: xi along y, eta along z
        for r in range(1, nribs + 1):
            i = r * bay
            for k in range(nh):
                for j in range(nc):
                    conn.append([node(i, j, k), node(i, j + 1, k), node(i, j, k + 1), node(i, j + 1, k + 1)])
                    panel.append(4 * nribs + r - 1)
        conn = np.array(conn, dtype=np.int64)
        X = np.array(coords)
        perm = first_appearance(conn, len(coords))
        self.conn = perm[conn].astype(np.int32)
        self.Xpts = np.empty_like(X)
        self.Xpts[perm] = X
        self.nnodes, self.nelems = len(coords), len(conn)
        self.panel_of = np.array(panel, dtype=np.int32)
        self.npanels = 5 * nribs
        # grid index of every node (new numbering)
        self.grid = np.empty((self.nnodes, 3), dtype=np.int64)
        for key, n in ids.items():
            self.grid[perm[n]] = key

    def bcs(self):
        nodes = np.where(self.grid[:, 0] == 0)[0].astype(np.int32)
        return nodes, np.full(nodes.size, 0x3F, dtype=np.int32)

    def root_nodes(self):
        """Node ids of the root station (a seed for the level ordering)."""
        return np.where(self.grid[:, 0] == 0)[0].astype(np.int32)

    def panel_thickness(self, seed, lo=0.002, hi=0.01):
        """Per-element thickness from one seeded uniform value per panel."""
        tp = np.random.default_rng(seed).uniform(lo, hi, self.npanels)
        return tp[self.panel_of]

    def load_case(self, c, ncases=16, tip_force=1.0e4, pressure=1.0e3):
        """Load case c of ncases: tip force tip_force cos(a) (N, +z, spread
        evenly over the tip rib's nodes) plus an upward pressure pressure
        sin(a) (Pa) on the bottom skin, lumped as a quarter of each element
        area per corner, with a = (c + 0.5) pi / (2 ncases)."""
        a = (c + 0.5) * np.pi / (2 * ncases)
        f = np.zeros((self.nnodes, 6))
        tip = np.where(self.grid[:, 0] == self.ns)[0]
        f[tip, 2] += tip_force * np.cos(a) / tip.size
        da = (self.span / self.ns) * (self.chord / self.nc)
        bottom = np.arange(self.nc * self.ns, 2 * self.nc * self.ns)
        for n in self.conn[bottom].ravel():
            f[n, 2] += pressure * np.sin(a) * da / 4.0
        return f.ravel()
