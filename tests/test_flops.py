"""bench.py's per-element flop constants equal the operation count of the
oracle's restatement of the reference element (oracle/flopcount)."""
import json
import os
import subprocess

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_flop_constants_match_counter():
    exe = os.path.join(ROOT, "oracle", "flopcount")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "flopcount"])
    counts = json.loads(subprocess.check_output([exe]).decode())
    assert counts == bench.FLOPS_PER_ELEMENT
