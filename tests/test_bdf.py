"""TACSMeshLoader (C++ host layer) reading Nastran bulk data, without a GPU:
a generated file with shuffled node and element ids, short / long / free
field GRIDs, a continued CQUAD4 and both SPC forms, read back through
lib/bdf_dump and compared with what was written; and the reference's own
fixture examples/cylinder-buckling/mech-cylinder.bdf (only where the
reference is mounted) against its stated geometry (R = 0.2, L = 0.4) and
boundary conditions."""
import os
import subprocess

import numpy as np
import pytest

from conftest import LIB_DIR
from shellgpu.bdf import bdf_text

DUMP = os.path.join(LIB_DIR, "bdf_dump")
REF_BDF = "/root/reference/examples/cylinder-buckling/mech-cylinder.bdf"


def _dump(path):
    out = subprocess.run([DUMP, path], check=True, capture_output=True, text=True).stdout.splitlines()
    X, E, B = {}, {}, {}
    for line in out:
        f = line.split()
        if f[0] == "X":
            X[int(f[1])] = [float(v) for v in f[2:5]]
        elif f[0] == "E":
            E[int(f[1])] = [int(v) for v in f[2:7]]
        elif f[0] == "B":
            B[int(f[1])] = (int(f[2]), [float(v) for v in f[3:9]])
    head = out[0].split()
    return int(head[1]), int(head[3]), X, E, B


@pytest.mark.skipif(not os.path.exists(DUMP), reason="host layer not built")
def test_generated_bdf_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    nx, ny = 3, 2
    xs, ys = np.meshgrid(np.linspace(0, 1, nx + 1), np.linspace(0, 0.5, ny + 1), indexing="ij")
    X = np.stack([xs, ys, 0.1 * xs * ys], -1).reshape(-1, 3)
    X = np.round(X + 0.01 * rng.standard_normal(X.shape), 6)
    nid = np.arange(X.shape[0]).reshape(nx + 1, ny + 1)
    conn = np.array([[nid[i, j], nid[i + 1, j], nid[i, j + 1], nid[i + 1, j + 1]]
                     for j in range(ny) for i in range(nx)])
    node_ids = rng.permutation(np.arange(10, 10 + 3 * X.shape[0], 3))
    elem_ids = rng.permutation(np.arange(1, conn.shape[0] + 1) * 7)
    spcs = [(0, "123456", 0.0), (3, "1", -1.5e-5), (3, "23", 0.0), (5, "246", 2.0)]
    text = bdf_text(X, conn, spcs, node_ids=node_ids, elem_ids=elem_ids, long_field=True).splitlines()
    # mix formats: a free-field GRID, a short-field GRID, a continued CQUAD4, a short SPC
    out = []
    i = 0
    while i < len(text):
        line = text[i]
        if line.startswith("GRID*") and int(line[8:24]) == node_ids[1]:
            x, y, z = X[1]
            out.append(f"GRID,{node_ids[1]},,{float(x)!r},{float(y)!r},{float(z)!r}")
            i += 2
            continue
        if line.startswith("GRID*") and int(line[8:24]) == node_ids[2]:
            x, y, z = X[2]
            out.append(f"{'GRID':<8}{node_ids[2]:>8d}{'':8}{x:>8.5f}{y:>8.5f}{z:>8.5f}")
            i += 2
            continue
        if line.startswith("CQUAD4") and int(line[8:16]) == elem_ids[0]:
            out.append(line[:40])
            out.append(" " * 8 + line[40:])
            i += 1
            continue
        if line.startswith("SPC*") and int(line[24:40]) == node_ids[5]:
            out.append(f"{'SPC':<8}{1:>8d}{node_ids[5]:>8d}{'246':>8}{'2.':>8}")
            i += 1
            continue
        out.append(line)
        i += 1
    path = tmp_path / "mesh.bdf"
    path.write_text("\n".join(out) + "\n")
    nn, ne, Xr, Er, Br = _dump(str(path))
    assert nn == X.shape[0] and ne == conn.shape[0]
    # nodes in file-id order, elements in element-id order
    order = np.argsort(node_ids)
    pos = np.empty_like(order)
    pos[order] = np.arange(len(order))
    for k in range(nn):
        src = order[k]
        tol = 1e-5 if src == 2 else 1e-14
        assert np.allclose(Xr[k], X[src], atol=tol, rtol=0), k
    eorder = np.argsort(elem_ids)
    for k in range(ne):
        e = eorder[k]
        assert Er[k][:4] == [int(pos[v]) for v in conn[e]], k
        assert Er[k][4] == 0
    assert Br[int(pos[0])][0] == 0x3F
    assert Br[int(pos[3])][0] == 0x7 and Br[int(pos[3])][1][0] == -1.5e-5
    assert Br[int(pos[5])][0] == (1 << 1) | (1 << 3) | (1 << 5) and Br[int(pos[5])][1][1] == 2.0


@pytest.mark.skipif(not (os.path.exists(DUMP) and os.path.exists(REF_BDF)), reason="reference fixture absent")
def test_reference_mech_cylinder_bdf():
    nn, ne, X, E, B = _dump(REF_BDF)
    assert (nn, ne) == (3300, 3200)
    P = np.array([X[k] for k in range(nn)])
    r = np.hypot(P[:, 1], P[:, 2])
    assert np.abs(r - 0.2).max() < 1e-9
    assert abs(P[:, 0].min()) < 1e-12 and abs(P[:, 0].max() - 0.4) < 1e-12
    # tensor order: corner 3 is opposite corner 0 (the diagonal is the longest side)
    for e in range(ne):
        c = E[e][:4]
        d01 = np.linalg.norm(P[c[1]] - P[c[0]])
        d02 = np.linalg.norm(P[c[2]] - P[c[0]])
        d03 = np.linalg.norm(P[c[3]] - P[c[0]])
        assert d03 > max(d01, d02)
    # both end rings: translations constrained, u1 = -1e-5 on one of them
    masks = [m for m, _ in B.values()]
    assert len(B) == 200 and set(masks) == {7}
    u1 = sorted(v[0] for _, v in B.values())
    assert u1.count(-1e-5) == 100 and u1.count(0.0) == 100
