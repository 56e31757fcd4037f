"""The stiffened box-beam generator: piece sizes, node sharing along the
junction lines, panels, root BCs and load-case totals (CPU only)."""
import numpy as np

from shellgpu.problems import BoxBeam


def test_full_size_counts():
    b = BoxBeam()
    # skins 2 x 400 x 80, spars 2 x 400 x 16, ribs 40 x 80 x 16
    assert b.nelems == 2 * 400 * 80 + 2 * 400 * 16 + 40 * 80 * 16 == 128000
    # skins (2 x 401 x 81) + spar interiors (2 x 401 x 15) + rib interiors (40 x 79 x 15)
    assert b.nnodes == 2 * 401 * 81 + 2 * 401 * 15 + 40 * 79 * 15 == 124392
    assert b.npanels == 200
    # panel sizes: skin bays 10 x 80, spar bays 10 x 16, ribs 80 x 16
    sizes = np.bincount(b.panel_of)
    assert list(np.unique(sizes)) == [160, 800, 1280]
    assert (sizes == 800).sum() == 80 and (sizes == 160).sum() == 80 and (sizes == 1280).sum() == 40


def test_small_beam_topology_and_loads():
    b = BoxBeam(20, 8, 2, 4, span=1.0, chord=0.2, height=0.05)
    conn = b.conn
    # every element: 4 distinct nodes forming a rectangle of the piece's size
    assert all(len(set(c)) == 4 for c in conn)
    X = b.Xpts[conn]
    area = np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1)
    assert np.allclose(area[:320], 0.05 * 0.025)      # skins
    assert np.allclose(area[320:400], 0.05 * 0.025)   # spars
    assert np.allclose(area[400:], 0.025 * 0.025)     # ribs
    # junction nodes: each skin-spar corner line is shared by 4 elements per
    # station (2 skin, 2 spar) away from ribs; rib-skin-spar corners by more
    count = np.bincount(conn.ravel(), minlength=b.nnodes)
    assert count.max() == 6
    # first-appearance numbering
    seen = []
    for n in conn.ravel():
        if n not in seen:
            seen.append(n)
    assert seen == list(range(b.nnodes))
    nodes, mask = b.bcs()
    assert np.all(b.Xpts[nodes, 0] == 0.0) and np.all(mask == 0x3F)
    assert nodes.size == 2 * 9 + 2 * 1  # root ring: 2 x 9 skin + 2 x 1 spar interior
    for c in (0, 7, 15):
        f = b.load_case(c).reshape(-1, 6)
        a = (c + 0.5) * np.pi / 32
        assert np.isclose(f[:, 2].sum(), 1e4 * np.cos(a) + 1e3 * np.sin(a) * 1.0 * 0.2)
        assert np.abs(f[:, [0, 1, 3, 4, 5]]).max() == 0.0
    t = b.panel_thickness(4000)
    assert t.min() >= 0.002 and t.max() <= 0.01
    assert all(np.unique(t[b.panel_of == p]).size == 1 for p in range(b.npanels))
