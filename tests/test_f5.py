"""The .f5 writer (TACSFH5File of the C++ host layer) against the reference's
own TACSFH5File compiled from source (oracle/_ref/fh5_ref_writer, built in
the development container only): the same zones give the same bytes."""
import os
import subprocess

import numpy as np
import pytest

from conftest import LIB_DIR
from f5read import read_f5

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_WRITER = os.path.join(ROOT, "oracle", "_ref", "fh5_ref_writer")


def build_own_writer(tmp_path):
    exe = str(tmp_path / "fh5_check")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "a2d-shells_amd", "host"),
                           "-I" + os.path.join(ROOT, "include"), "-o", exe, os.path.join(HERE, "fh5_check.cpp"),
                           "-L" + LIB_DIR, "-ltacsgpu", "-lshellgpu", "-Wl,-rpath," + LIB_DIR])
    return exe


def test_writer_layout(tmp_path):
    exe = build_own_writer(tmp_path)
    out = str(tmp_path / "own.f5")
    subprocess.check_call([exe, out])
    comps, zones = read_f5(out)
    assert comps == ["Component 0", "Skin panels"]
    assert [z["name"] for z in zones] == ["components", "continuous data t=0.0000000000e+00",
                                          "element data t=0.0000000000e+00"]
    assert zones[0]["data"].ravel().tolist() == [0, 1, 1, 0, 1]
    assert np.array_equal(zones[1]["data"].ravel(), 0.25 * np.arange(27, dtype=np.float32) - 1.0)
    assert zones[1]["vars"] == ["X", "Y", "Z", "u", "v", "w", "rotx", "roty", "rotz"]
    assert zones[2]["data"].dtype == np.float64 and zones[2]["data"].shape == (3, 2)


@pytest.mark.skipif(not os.path.exists(REF_WRITER), reason="reference writer not built (make -C oracle ref)")
def test_writer_bytes_match_reference(tmp_path):
    exe = build_own_writer(tmp_path)
    own, ref = str(tmp_path / "own.f5"), str(tmp_path / "ref.f5")
    subprocess.check_call([exe, own])
    subprocess.check_call([REF_WRITER, ref])
    assert open(own, "rb").read() == open(ref, "rb").read()
