"""The traversal's FMA slab form against the reference's slab arithmetic
(rayAabbIntersection, udpt.cl:392-431), on the CPU: tests/slab_check.cpp
runs the device header yune_amd/csrc/yn_slab.h built with the kernels'
floating-point rules on random and adversarial rays and boxes.  Every box
the reference's arithmetic enters must be entered (and not culled) by the
FMA form; among boxes shaped like a BVH's, at most 1 in 1000 more."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def slab_check(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("slab") / "slab_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    "-I", os.path.join(ROOT, "yune_amd", "csrc"), os.path.join(ROOT, "tests", "slab_check.cpp"),
                    "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_fma_slab_form_enters_every_box_the_reference_enters(slab_check, seed):
    r = subprocess.run([slab_check, "4000000", str(seed)], capture_output=True, text=True, timeout=120)
    out = r.stdout
    m = re.search(r"tested (\d+) boxes, (\d+) entered by the reference, (\d+) more by the FMA form "
                  r"\((\d+) of (\d+) like a BVH's\), (\d+) missed; (\d+) cull tests, (\d+) missed", out)
    assert m, out
    tested, ref, extra, bvh_extra, bvh, missed, culls, cull_missed = map(int, m.groups())
    assert tested > 3_000_000 and ref > 500_000 and culls > 1_000_000
    assert missed == 0 and cull_missed == 0, out
    assert bvh > 100_000 and bvh_extra * 1000 <= bvh, out
    assert r.returncode == 0, out
