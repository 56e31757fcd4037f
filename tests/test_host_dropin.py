"""Drop-in check of the C++ host classes: the reference's own buckling
driver (examples/cylinder-buckling/mechBuckling.cpp), with its include lines
replaced by TACSShellGPU.h and `using namespace shellgpu;` and nothing else
changed, compiles and links against libtacsgpu.so (oracle/Makefile target
_ref/mech_dropin; the edited copy stays in oracle/_ref/).  Runs where the
reference sources are present (this check reads them)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DRIVER = "/root/reference/examples/cylinder-buckling/mechBuckling.cpp"


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="reference sources absent")
def test_reference_driver_builds_against_host_classes():
    lib = os.path.join(ROOT, "a2d-shells_amd", "lib", "libtacsgpu.so")
    if not os.path.exists(lib):
        pytest.skip("libtacsgpu.so not built")
    out = subprocess.run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "_ref/mech_dropin"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    src = open(os.path.join(ROOT, "oracle", "_ref", "mechBuckling_dropin.cpp")).read().splitlines()
    ref = open(REF_DRIVER).read().splitlines()
    # the only changes: two lines in front, the reference's own includes gone
    assert src[:2] == ['#include "TACSShellGPU.h"', "using namespace shellgpu;"]
    assert src[2:] == [line for line in ref if not line.startswith('#include "')]
    assert os.access(os.path.join(ROOT, "oracle", "_ref", "mech_dropin"), os.X_OK)
