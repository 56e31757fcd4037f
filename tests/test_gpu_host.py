"""The C++ host layer end to end on the GPU: the port of the reference's
mechBuckling driver (lib/mech_buckling) reads a BDF of a cylinder with the
reference fixture's geometry and boundary conditions (R = 0.2, L = 0.4,
translations of both end rings constrained, u1 = -1e-5 prescribed on one
ring), written with shuffled node and element ids, and its buckling loads
are compared with the oracle's on the same mesh."""
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as O
from conftest import LIB_DIR
from shellgpu.bdf import write_bdf
from shellgpu.problems import Cylinder

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(LIB_DIR, "mech_buckling")


def test_mech_buckling_driver_matches_oracle(tmp_path):
    cyl = Cylinder(10, 20, R=0.2, L=0.4)
    t, sigma, nev, nlanczos = 0.002, 10.0, 8, 120
    ring0, ring1 = cyl.ring_nodes(0), cyl.ring_nodes(cyl.na)
    spcs = [(int(n), "123", 0.0) for n in ring0]
    spcs += [(int(n), "1", -1e-5) for n in ring1] + [(int(n), "23", 0.0) for n in ring1]
    rng = np.random.default_rng(9)
    node_ids = rng.permutation(np.arange(1, cyl.nnodes + 1) * 2)
    # element ids increasing in element order: first-appearance numbering
    # then reproduces the generator's node order
    elem_ids = np.arange(1, cyl.nelems + 1) * 3
    path = str(tmp_path / "cyl.bdf")
    write_bdf(path, cyl.Xpts, cyl.conn, spcs, node_ids=node_ids, elem_ids=elem_ids)
    # the driver runs as a child process: its own HIP context
    out = subprocess.run([DRIVER, path, str(t), str(sigma), str(nev), str(nlanczos)], capture_output=True,
                         text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    lam = np.array([float(m.group(2)) for m in re.finditer(r"lambda\[(\d+)\] = (\S+)", out.stdout)])
    assert lam.size == nev

    # the BDF's coordinates are written with 9 significant digits
    Xf = np.array([[float(f"{v:.9e}") for v in row] for row in cyl.Xpts])
    bnodes = np.concatenate([ring0, ring1]).astype(np.int32)
    masks = np.full(bnodes.size, 7, np.int32)
    vals = np.zeros((bnodes.size, 6))
    vals[ring0.size:, 0] = -1e-5
    ref = O.linear_buckling(Xf, cyl.conn, np.full(cyl.nelems, t), (bnodes, masks, vals.ravel()),
                            np.zeros(6 * cyl.nnodes), sigma, nev=nev, max_lanczos=nlanczos,
                            mat=O.material(E=72e9, nu=0.33, rho=2718.0, cte=10e-6))[0]
    print("driver", lam[:5], "oracle", ref[:5])
    assert np.abs(lam[:5] - ref[:5]).max() <= 1e-8 * np.abs(ref[:5]).max()


FREQ_DRIVER = os.path.join(LIB_DIR, "mech_frequency")


def test_mech_frequency_driver_matches_dense_oracle(tmp_path):
    """TACSFrequencyAnalysis through the C++ classes on a BDF cylinder with
    both end rings clamped: omega^2 against a dense solve of the oracle's K
    and M (M x = mu K x on the free dofs, lambda = 1 / mu), and the element
    sum of d(lambda_0)/dt against a central difference of the dense value
    under a uniform thickness change, and X . d(lambda_0)/dX against one
    under a uniform scaling of the mesh."""
    import scipy.linalg as sla

    cyl = Cylinder(8, 16, R=0.2, L=0.4)
    t, nev = 0.002, 6
    ring0, ring1 = cyl.ring_nodes(0), cyl.ring_nodes(cyl.na)
    spcs = [(int(n), "123456", 0.0) for n in np.concatenate([ring0, ring1])]
    elem_ids = np.arange(1, cyl.nelems + 1)
    path = str(tmp_path / "cyl.bdf")
    write_bdf(path, cyl.Xpts, cyl.conn, spcs, node_ids=np.arange(1, cyl.nnodes + 1), elem_ids=elem_ids)
    out = subprocess.run([FREQ_DRIVER, path, str(t), "0", str(nev), "100"], capture_output=True, text=True,
                         timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    lam = np.array([float(m.group(2)) for m in re.finditer(r"lambda\[(\d+)\] = (\S+)", out.stdout)])
    dsum = float(re.search(r"sum d\(lambda_0\)/dt = (\S+)", out.stdout).group(1))
    assert lam.size == nev

    Xf = np.array([[float(f"{v:.9e}") for v in row] for row in cyl.Xpts])
    m = O.material(E=72e9, nu=0.33, rho=2718.0, cte=10e-6)
    rowp, cols = O.node_csr(cyl.nnodes, cyl.conn)
    free = np.ones(6 * cyl.nnodes, bool)
    for n in np.concatenate([ring0, ring1]):
        free[6 * n:6 * n + 6] = False

    def dense(tt, X=Xf):
        th = np.full(cyl.nelems, tt)
        K = O.bsr_to_dense(rowp, cols, O.assemble_mat(O.KIND_K, 0, m, X, cyl.conn, th, rowp, cols))
        M = O.bsr_to_dense(rowp, cols, O.assemble_mat(O.KIND_M, 0, m, X, cyl.conn, th, rowp, cols))
        mu = sla.eigh(M[np.ix_(free, free)], K[np.ix_(free, free)], eigvals_only=True)
        return np.sort(1.0 / mu[mu > 0])[:nev]

    ref = dense(t)
    print("driver", lam, "dense", ref)
    assert np.abs(lam[:5] - ref[:5]).max() <= 1e-8 * np.abs(ref[:5]).max()
    # d lambda_0 / d(uniform t) = sum_e d lambda_0 / d t_e; lambda_0 is a
    # degenerate pair member on the perfect cylinder, but a uniform thickness
    # change keeps the symmetry, so the pair moves together
    h = 1e-6 * t
    fd = (dense(t + h)[0] - dense(t - h)[0]) / (2 * h)
    assert abs(dsum - fd) <= 1e-6 * abs(fd)
    # X . d(lambda_0)/dX: the derivative under a uniform scaling of the mesh,
    # which also keeps the pair together
    dscale = float(re.search(r"X \. d\(lambda_0\)/dX = (\S+)", out.stdout).group(1))
    s = 1e-6
    fd = (dense(t, Xf * (1 + s))[0] - dense(t, Xf * (1 - s))[0]) / (2 * s)
    print("scaling derivative", dscale, fd)
    assert abs(dscale - fd) <= 1e-6 * abs(fd)


GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reference_bdf_buckling_matches_oracle_golden(tmp_path):
    """The reference's own example mesh (examples/cylinder-buckling/
    mech-cylinder.bdf, kept compressed under tests/golden/): the C++ driver's
    loads against the oracle's on the same file (tests/golden/
    make_bdf_golden.py), the first 5 within 1e-8."""
    import gzip
    import json

    gold = json.load(open(os.path.join(GOLDEN_DIR, "mech_cylinder_buckling.json")))
    path = str(tmp_path / "mech-cylinder.bdf")
    with gzip.open(os.path.join(GOLDEN_DIR, "mech-cylinder.bdf.gz"), "rb") as f:
        open(path, "wb").write(f.read())
    out = subprocess.run([DRIVER, path, str(gold["thickness"]), str(gold["sigma"]), str(gold["nev"]),
                          str(gold["max_lanczos"])], capture_output=True, text=True, timeout=180)
    print(out.stdout[-2500:], out.stderr[-2000:])
    assert out.returncode == 0
    assert f"nodes {gold['nodes']} elements {gold['elements']}" in out.stdout
    lam = np.array([float(m.group(2)) for m in re.finditer(r"lambda\[(\d+)\] = (\S+)", out.stdout)])
    ref = np.array(gold["lambda"])
    print("driver", lam[:5], "oracle", ref[:5])
    assert np.abs(lam[:5] - ref[:5]).max() <= 1e-8 * np.abs(ref[:5]).max()
