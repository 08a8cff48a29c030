"""GPU participating media: closed forms and parity with the oracle, through the C ABI.

* Beer-Lambert: camera paths through a purely absorbing slab (delta tracking) and NEE shadow rays
  through one (ratio tracking) against exp(-sigma_a d) per spectral bin.
* Per-pixel 3-sigma parity with the oracle's converged renders of a glass sphere holding a
  chromatic HG medium, of a fog box with nested media and the camera inside (medium tracker,
  priorities, volume boundaries), and of the exponential atmosphere (FP64 shells and optical depth).
"""
import numpy as np
import pytest

from naivetrace_amd import abi, configs
from naivetrace_amd.renderer import GpuRenderer
from test_oracle import absorber_expected, point_plane_expected
from test_gpu_parity import render_parity

pytestmark = pytest.mark.gpu


def _mean_se(scene, w, h, spp, depth, mode=abi.SAMPLE_BOTH, seed=13):
    p = abi.default_params(w, h, max_depth=depth, seed=seed)
    p.sample_mode = mode
    with GpuRenderer(scene, p) as r:
        r.render_spp(spp)
        rgb, sq, n = r.framebuffer()
    assert np.all(n == spp)
    N = w * h * spp
    mean = rgb.sum(axis=(0, 1)) / N
    var = sq.sum(axis=(0, 1)) / N - mean ** 2
    return mean, np.sqrt(var / N)


@pytest.mark.parametrize("mode", [abi.SAMPLE_BOTH, abi.SAMPLE_BSDF], ids=["both", "bsdf"])
def test_gpu_absorber_transmittance(gpu_lib, mode):
    s = configs.absorber_slab(16, 16)
    mean, se = _mean_se(s, 16, 16, 4096, 4, mode)
    expected = absorber_expected(s, 0.5)
    print("absorber", mean, "+-", se, "expected", expected, flush=True)
    # FP32 transmittance: 4 standard errors plus 2e-3 relative.
    assert np.all(np.abs(mean - expected) < 4 * se + 2e-3 * expected)


def test_gpu_tabulated_absorber_shell(gpu_lib):
    # Altitude-constant tabulated absorber inside a sphere of radius 10: L exp(-sigma_a * 10) per bin.
    s = configs.tabulated_absorber_shell(16, 16)
    mean, se = _mean_se(s, 16, 16, 4096, 4)
    expected = absorber_expected(s, s.shell_radius)
    print("tabulated absorber", mean, "+-", se, "expected", expected, flush=True)
    assert np.all(np.abs(mean - expected) < 4 * se + 2e-3 * expected)


def test_gpu_nee_transmission(gpu_lib):
    s = configs.point_lit_plane_behind_absorber(16, 16)
    mean, se = _mean_se(s, 16, 16, 1024, 2)
    expected = point_plane_expected(s)
    print("nee transmission", mean, "+-", se, "expected", expected, flush=True)
    assert np.all(np.abs(mean - expected) < 4 * se + 2e-3 * expected)


def test_gpu_medium_sphere_parity(gpu_lib):
    render_parity("medium_sphere32", builder=configs.medium_sphere)


def test_gpu_fog_box_parity(gpu_lib):
    render_parity("fog_box32", builder=configs.fog_box)


def test_gpu_atmosphere_parity(gpu_lib):
    render_parity("atmosphere32", builder=configs.atmosphere)


@pytest.mark.parametrize("variant", ["empty", "boundary"])
@pytest.mark.parametrize("mode", [abi.SAMPLE_BOTH, abi.SAMPLE_BSDF], ids=["both", "bsdf"])
def test_gpu_fog_furnace(gpu_lib, variant, mode):
    """Non-absorbing fog in the emitting furnace keeps the equilibrium radiance L_e / (1 - rho)."""
    from test_oracle import fog_furnace_scene
    from naivetrace_amd.scene import flat, spectrum_to_rgb
    rho = 0.5
    s = fog_furnace_scene(rho, variant)
    p = abi.default_params(32, 32, max_depth=80, seed=21)
    p.mis_mode = abi.MIS_CONSISTENT
    p.sample_mode = mode
    with GpuRenderer(s, p) as r:
        r.render_spp(256)
        rgb, sq, n = r.framebuffer()
    N = 32 * 32 * 256
    mean = rgb.sum(axis=(0, 1)) / N
    se = np.sqrt((sq.sum(axis=(0, 1)) / N - mean ** 2) / N)
    expected = spectrum_to_rgb(flat(1.0 / (1 - rho)))
    print(variant, mode, mean, "+-", se, "expected", expected, flush=True)
    assert np.all(np.abs(mean - expected) < 4 * se + 2e-3 * np.abs(expected))
