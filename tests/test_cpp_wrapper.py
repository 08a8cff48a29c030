"""The C++ RenderingAlgo mirror (include/ttgpu_algo.hpp) over the C ABI, compiled with g++
against the in-tree libttgpu.so: exceptions for invalid scenes (CPU), and a closed-form render
through the C++ interface (GPU: plane under a dome, radiance = albedo x dome radiance)."""
import os
import subprocess

import numpy as np
import pytest

from naivetrace_amd import abi
from naivetrace_amd.scene import flat, spectrum_to_rgb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "naivetrace_amd")


@pytest.fixture(scope="module")
def demo(tmp_path_factory):
    abi.load_library()  # raises if the library is missing
    exe = str(tmp_path_factory.mktemp("cpp") / "wrapper_demo")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "wrapper_demo.cpp"),
                    "-L", LIBDIR, "-lttgpu", f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True)
    return exe


def test_cpp_wrapper_invalid_scene_throws(demo):
    r = subprocess.run([demo, "invalid"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "invalid_argument" in r.stdout and "material" in r.stdout


@pytest.mark.gpu
def test_cpp_wrapper_render_closed_form(demo):
    r = subprocess.run([demo, "render", "256"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    parts = r.stdout.split()
    mean = np.array([float(x) for x in parts[parts.index("mean") + 1:parts.index("mean") + 4]])
    expected = spectrum_to_rgb(flat(0.6 * 1.5))
    print(r.stdout.strip(), "expected", expected, flush=True)
    # 1 + 2 + 4 + ... passes up to 256 spp; 32 x 32 x 256 paths of a one-bounce estimator.
    assert "passes 9 spp 256" in r.stdout
    np.testing.assert_allclose(mean, expected, rtol=1e-2)
