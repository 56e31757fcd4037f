"""Element output data (getOutputData: strains, stresses, failure and
thickness at the element nodes) on the GPU against the oracle, both strain
models; and the .f5 files the C++ buckling driver writes per mode
(TACSToFH5): zones, names, connectivity, nodes, mode shapes, and element data
consistent with the oracle evaluated on the file's own (float32) nodes and
displacements."""
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as O
import shellgpu as sg
from conftest import LIB_DIR
from f5read import read_f5
from helpers import random_mesh
from shellgpu.bdf import write_bdf
from shellgpu.problems import Cylinder

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model", [sg.LINEAR, sg.NONLINEAR])
def test_element_output_matches_oracle(model):
    conn, X = random_mesh(5, 4, 7)
    ne, nn = conn.shape[0], X.shape[0]
    rng = np.random.default_rng(8)
    thick = rng.uniform(0.0015, 0.0025, ne)
    u = 2e-3 * rng.standard_normal(6 * nn)
    ys = 3e6
    ctx = sg.Context(0)
    ctx.set_mesh(conn, X)
    ctx.set_material(70e9, 0.3, 2700.0, 0.0, 10.0)
    ctx.set_designs(thick)
    out = ctx.element_output(u, model=model, ys=ys)[0]
    ctx.close()
    m = O.material()
    ref = np.stack([O.elem_output(model, m, thick[e], ys, X[conn[e]].ravel(),
                                  u.reshape(-1, 6)[conn[e]].ravel()) for e in range(ne)])
    for lo, hi in [(0, 9), (9, 18), (18, 19)]:  # strains, stresses, failure0
        err = np.abs(out[..., lo:hi] - ref[..., lo:hi]).max() / np.abs(ref[..., lo:hi]).max()
        print(lo, hi, err)
        assert err <= 1e-10
    assert np.array_equal(out[..., 24], ref[..., 24])  # dv1 = t
    assert np.abs(out[..., 19:24]).max() == 0.0 and np.abs(out[..., 25:]).max() == 0.0


def test_buckling_driver_writes_mode_files(tmp_path):
    cyl = Cylinder(8, 16, R=0.2, L=0.4)
    ring0, ring1 = cyl.ring_nodes(0), cyl.ring_nodes(cyl.na)
    spcs = [(int(n), "123", 0.0) for n in ring0]
    spcs += [(int(n), "1", -1e-5) for n in ring1] + [(int(n), "23", 0.0) for n in ring1]
    path = str(tmp_path / "cyl.bdf")
    write_bdf(path, cyl.Xpts, cyl.conn, spcs, node_ids=np.arange(1, cyl.nnodes + 1),
              elem_ids=np.arange(1, cyl.nelems + 1))
    prefix = str(tmp_path / "mode")
    t = 0.002
    out = subprocess.run([os.path.join(LIB_DIR, "mech_buckling"), path, str(t), "10", "3", "80", prefix],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout[-1500:], out.stderr[-1500:])
    assert out.returncode == 0
    assert len(re.findall(r"lambda\[\d+\]", out.stdout)) == 3
    m = O.material(E=72e9, nu=0.33, rho=2718.0, cte=10e-6)
    for mode in range(3):
        comps, zones = read_f5(f"{prefix}{mode}.f5")
        assert comps == ["Component 0"]
        names = [z["name"] for z in zones]
        assert names == ["components", "ltypes", "ptr", "connectivity", "continuous data t=0.0000000000e+00",
                         "element data t=0.0000000000e+00"]
        z = {zz["name"]: zz for zz in zones}
        ne, nn = cyl.nelems, cyl.nnodes
        assert np.all(z["components"]["data"] == 0) and np.all(z["ltypes"]["data"] == 8)
        assert np.array_equal(z["ptr"]["data"].ravel(), 4 * np.arange(ne + 1))
        conn = z["connectivity"]["data"].reshape(ne, 4)
        cont = z["continuous data t=0.0000000000e+00"]
        assert cont["vars"] == ["X", "Y", "Z", "u", "v", "w", "rotx", "roty", "rotz"]
        Xf = cont["data"][:, :3].astype(np.float64)
        uf = cont["data"][:, 3:].astype(np.float64)
        assert Xf.shape == (nn, 3) and np.abs(uf).max() > 0
        # the file's mesh is the driver's (renumbered) mesh: same element
        # areas as the generator's
        A = np.linalg.norm(np.cross(Xf[conn[:, 1]] - Xf[conn[:, 0]], Xf[conn[:, 2]] - Xf[conn[:, 0]]), axis=1)
        A0 = np.linalg.norm(np.cross(cyl.Xpts[cyl.conn[:, 1]] - cyl.Xpts[cyl.conn[:, 0]],
                                     cyl.Xpts[cyl.conn[:, 2]] - cyl.Xpts[cyl.conn[:, 0]]), axis=1)
        assert np.allclose(np.sort(A), np.sort(A0), rtol=1e-5)
        el = z["element data t=0.0000000000e+00"]
        assert el["vars"][:3] == ["ex0", "ey0", "exy0"] and el["vars"][-1] == "dv7" and len(el["vars"]) == 31
        data = el["data"].reshape(ne, 4, 31).astype(np.float64)
        ref = np.stack([O.elem_output(0, m, t, 1e11, Xf[conn[e]].ravel(), uf[conn[e]].ravel())
                        for e in range(ne)])
        for lo, hi in [(0, 9), (9, 18)]:
            err = np.abs(data[..., lo:hi] - ref[..., lo:hi]).max() / np.abs(ref[..., lo:hi]).max()
            assert err <= 1e-3, (mode, lo, err)
        assert np.allclose(data[..., 24], t, rtol=1e-6)
