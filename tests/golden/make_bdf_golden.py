"""Write tests/golden/mech_cylinder_buckling.json: the oracle's first
buckling loads of the reference's example mesh.

tests/golden/mech-cylinder.bdf.gz is the reference's input data file
examples/cylinder-buckling/mech-cylinder.bdf (3,300 GRID*, 3,200 CQUAD4,
SPC* with the u1 = -1e-5 end shortening), compressed.  This script reads it
with the C++ host layer's TACSMeshLoader (lib/bdf_dump, CPU only), runs the
oracle's linear buckling with the reference driver's material and thickness
(examples/cylinder-buckling/mechBuckling.cpp: aluminium, t = 0.010) and a
shift just below the first load (sigma = 1100; the driver's sigma = 10 leaves
the fifth and sixth loads unconverged after 150 Lanczos steps) and stores
the loads; the GPU test runs lib/mech_buckling on the same file with the
same settings and compares.

    python tests/golden/make_bdf_golden.py
"""
import gzip
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402

T, SIGMA, NEV, NLANCZOS = 0.010, 1100.0, 8, 150


def read_dump(path):
    out = subprocess.run([os.path.join(ROOT, "a2d-shells_amd", "lib", "bdf_dump"), path], check=True,
                         capture_output=True, text=True).stdout.splitlines()
    X, conn, bn, bm, bv = [], [], [], [], []
    for line in out:
        f = line.split()
        if f[0] == "X":
            X.append([float(v) for v in f[2:5]])
        elif f[0] == "E":
            conn.append([int(v) for v in f[2:6]])
        elif f[0] == "B":
            bn.append(int(f[1]))
            bm.append(int(f[2]))
            bv.append([float(v) for v in f[3:9]])
    return np.array(X), np.array(conn, np.int32), (np.array(bn, np.int32), np.array(bm, np.int32),
                                                  np.array(bv).ravel())


def main():
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "mech-cylinder.bdf")
        with gzip.open(os.path.join(HERE, "mech-cylinder.bdf.gz"), "rb") as f:
            open(path, "wb").write(f.read())
        X, conn, bcs = read_dump(path)
    m = O.material(E=72e9, nu=0.33, rho=2718.0, cte=10e-6)
    t0 = time.time()
    lam, phi, upath, resid, ns = O.linear_buckling(X, conn, np.full(conn.shape[0], T), bcs, np.zeros(X.size * 2),
                                                   SIGMA, nev=NEV, max_lanczos=NLANCZOS, mat=m)
    print("oracle", time.time() - t0, "s", lam, resid)
    json.dump({"thickness": T, "sigma": SIGMA, "nev": NEV, "max_lanczos": NLANCZOS, "lambda": lam.tolist(),
               "ritz_residual": resid.tolist(), "nodes": int(X.shape[0]), "elements": int(conn.shape[0])},
              open(os.path.join(HERE, "mech_cylinder_buckling.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
