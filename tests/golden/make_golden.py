"""Write tests/golden/element_golden.npz from the reference's own element code.

Run in the development container only (it needs /root/reference):

    make -C oracle ref
    python tests/golden/make_golden.py

The reference element sources are compiled from /root/reference by
oracle/Makefile (target ``ref``) into oracle/_ref/ref_driver; this script
builds the input records, runs that driver once and stores inputs and outputs
as a small fixture.  Record layout is documented in oracle/ref_driver.cpp.

Cases (all FP64, seeded):
  * random distorted quads, small and large displacement, T = 0
  * random distorted quads with T != 0 (thermal strain path)
  * elements cut from the 20x40 reference cylinder (R = 1, L = 2) with random
    element states
  * a flat unit plate element
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

E, NU, RHO, CTE = 70e9, 0.3, 2700.0, 1e-5


def cylinder_element(i, j, na=20, nc=40, R=1.0, L=2.0):
    def node(a, b):
        th = 2 * np.pi * (b % nc) / nc
        return [a * L / na, R * np.cos(th), R * np.sin(th)]

    return np.array([node(i, j), node(i + 1, j), node(i, j + 1), node(i + 1, j + 1)])


def make_records():
    rng = np.random.default_rng(20261202)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float)
    recs, tags = [], []

    def add(X, q, psi, t, temp, tag):
        recs.append(np.concatenate([X.ravel(), q, psi, [t, E, NU, RHO, CTE, temp]]))
        tags.append(tag)

    for k in range(4):
        X = 0.5 * base + 0.08 * rng.standard_normal((4, 3))
        q = (1e-4 if k < 2 else 2e-2) * rng.standard_normal(24)
        add(X, q, rng.standard_normal(24), rng.uniform(0.0015, 0.0025), 0.0,
            "random_small" if k < 2 else "random_large")
    for k in range(2):
        X = 0.5 * base + 0.08 * rng.standard_normal((4, 3))
        q = 1e-4 * rng.standard_normal(24)
        add(X, q, rng.standard_normal(24), 0.002, 10.0 * (k + 1), "thermal")
    for (i, j) in [(0, 0), (7, 13), (19, 39), (10, 21)]:
        X = cylinder_element(i, j)
        q = 1e-6 * rng.standard_normal(24)
        add(X, q, rng.standard_normal(24), rng.uniform(0.0015, 0.0025), 0.0,
            "cylinder")
    add(base.copy(), 1e-3 * rng.standard_normal(24), rng.standard_normal(24),
        0.01, 0.0, "flat_plate")
    return np.array(recs), tags


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -C oracle ref")
    recs, tags = make_records()
    tmp_in = os.path.join(ROOT, "oracle", "_ref", "golden_in.bin")
    tmp_out = os.path.join(ROOT, "oracle", "_ref", "golden_out.bin")
    recs.tofile(tmp_in)
    subprocess.check_call([DRIVER, tmp_in, tmp_out])
    out = np.fromfile(tmp_out).reshape(len(recs), -1)
    np.savez_compressed(
        os.path.join(HERE, "element_golden.npz"),
        Xpts=recs[:, 0:12], vars=recs[:, 12:36], psi=recs[:, 36:60],
        thickness=recs[:, 60], E=recs[:, 61], nu=recs[:, 62], rho=recs[:, 63],
        cte=recs[:, 64], temperature=recs[:, 65], tags=np.array(tags),
        res_lin=out[:, 0:24], res_nl=out[:, 24:48],
        K_lin=out[:, 48:624].reshape(-1, 24, 24),
        K_nl=out[:, 624:1200].reshape(-1, 24, 24),
        G=out[:, 1200:1776].reshape(-1, 24, 24),
        adj_lin=out[:, 1776], dRdt_lin=out[:, 1777], dRdt_nl=out[:, 1778],
        energy_lin=out[:, 1779], M=out[:, 1780:2356].reshape(-1, 24, 24),
        output_lin=out[:, 2356:2516].reshape(-1, 4, 40), ys=1e11)
    print("wrote", len(recs), "records")


if __name__ == "__main__":
    main()
