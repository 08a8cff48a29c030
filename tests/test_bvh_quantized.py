"""The quantized surface BVH (8-bit child planes, 64-byte nodes) on the host: the traversal source
of the kernels (naivetrace_amd/csrc/trace.hpp) compiled for the CPU by hipcc
(tests/cpp/bvh_quantized_check.cpp).  Decoded boxes contain the true ones; near the origin both
the quantized and the full-precision tree give exactly the brute-force closest hit and any-hit
answer.  Far from it (x = -1e6, unit triangles, float spacing 1/16) the slab test's cancellation
(as in the reference's non-robust Embree mode) loses about 0.3 % of the hits in both trees; the
quantized tree, whose boxes are slightly larger, loses no more than the full one.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "naivetrace_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    from naivetrace_amd import build
    exe = str(tmp_path_factory.mktemp("qbvh") / "bvh_quantized_check")
    subprocess.run([build.hipcc(), "-O2", "-std=c++17", "-x", "hip", f"--offload-arch={build.ARCH}", "-Wno-unused-result",
                    os.path.join(ROOT, "tests", "cpp", "bvh_quantized_check.cpp"), os.path.join(CSRC, "bvh_build.cpp"),
                    "-o", exe, "-lpthread"], check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("grid,offset,exact", [(60, 0.0, True), (80, 5000.0, True), (8, -1e6, False)])
def test_quantized_bvh_matches_brute_force(checker, grid, offset, exact):
    r = subprocess.run([checker, str(grid), str(offset), "40000"], capture_output=True, text=True, timeout=300)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    val = lambda k: int(f[f.index(k) + 1])
    assert val("box_errors") == 0 and val("field_errors") == 0
    assert val("hits") > 500
    if exact:
        for k in ("quantized_closest_diff", "quantized_anyhit_diff", "full_closest_diff", "full_anyhit_diff"):
            assert val(k) == 0, k
    else:
        assert val("quantized_closest_diff") <= val("full_closest_diff")
        assert val("quantized_anyhit_diff") <= val("full_anyhit_diff")
        assert val("quantized_closest_diff") < 0.01 * val("rays")
