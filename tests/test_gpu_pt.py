"""pt (the forward path tracer of the random-walk framework, the reference's
default --algo) on the GPU, through the C ABI.

* Parity with the oracle's renders (tests/golden/*_pt_oracle.npz): the Cornell box
  with both techniques and with light connections only, and a wide view of the box
  with a visible point light and a dome (the reference's truncating pixel lookup).
  Light connections land in other pixels than their camera sample, so both sides'
  standard errors come from independent batches.
* Closed forms: furnace L_e / (1 - rho) in all sample modes; an emitter filling
  the view gives L_e in every pixel; with the reference's pixel lookup the first
  column and row collect the extra band.
"""
import os

import numpy as np
import pytest

from naivetrace_amd import abi, configs
from naivetrace_amd.renderer import GpuRenderer, TTGpuError
from naivetrace_amd.scene import flat, spectrum_to_rgb
from test_gpu_parity import GOLDEN, _zscores
from test_oracle import fog_furnace_scene, furnace_scene
from test_pt import edge_pattern, emitter_wall_scene
from golden.make_oracle_images import PT_FIXTURES

pytestmark = pytest.mark.gpu


def _params(w, h, depth, mode, mis, seed=11):
    p = abi.default_params(w, h, max_depth=depth, seed=seed, algo=abi.ALGO_PT)
    p.sample_mode = mode
    p.mis_mode = mis
    return p


def gpu_batches(scene, params, batches, spp):
    """Per-pixel mean and standard error of `batches` independent renders of `spp` samples."""
    means = []
    with GpuRenderer(scene, params) as r:
        for _ in range(batches):
            r.render_spp(spp)
            rgb, _, n = r.framebuffer()
            assert np.all(n == spp)
            means.append(rgb / spp)
            r.reset()
    means = np.array(means)
    return means.mean(axis=0), means.std(axis=0, ddof=1) / np.sqrt(batches)


@pytest.mark.parametrize("name", list(PT_FIXTURES))
def test_gpu_pt_parity(gpu_lib, name):
    build, w, h, batches, spp, depth, mis, mode = PT_FIXTURES[name]
    f = np.load(os.path.join(GOLDEN, f"{name}_oracle.npz"))
    assert int(f["mis_mode"]) == mis and int(f["sample_mode"]) == mode and int(f["algo"]) == abi.ALGO_PT
    print(f"pt parity {name}: {batches} x {spp} spp", flush=True)
    mean, se = gpu_batches(build(w, h), _params(w, h, depth, mode, mis), batches, spp)
    z = _zscores(mean, se, f["mean"], f["stderr"])
    frac3 = float(np.mean(np.abs(z) > 3))
    print(f"{name}: mean z {z.mean():+.4f}, |z|>3 fraction {frac3:.4f}, max |z| {np.abs(z).max():.2f}", flush=True)
    # 8 batches: the z-scores follow a t distribution with 7 degrees of freedom on each side
    # (1.5 % beyond |3| for a t_7 alone), so allow 4 %.
    assert frac3 < 0.04
    assert abs(z.mean()) < 5.0 / np.sqrt(z.size) + 0.03
    img, img_ref = mean.mean(axis=(0, 1)), f["mean"].mean(axis=(0, 1))
    img_se = np.sqrt((se ** 2).sum(axis=(0, 1)) + (f["stderr"] ** 2).sum(axis=(0, 1))) / (w * h)
    print(f"{name}: image mean {img} oracle {img_ref} +- {img_se}", flush=True)
    assert np.all(np.abs(img - img_ref) < 4 * img_se + 1e-3 * np.abs(img_ref))


@pytest.mark.parametrize("mode", [abi.SAMPLE_BOTH, abi.SAMPLE_BSDF, abi.SAMPLE_LIGHTS])
def test_gpu_pt_furnace_closed_form(gpu_lib, mode):
    rho = 0.5
    mean, se = gpu_batches(furnace_scene(rho), _params(8, 8, 60, mode, abi.MIS_CONSISTENT), 8, 256)
    img, img_se = mean.mean(axis=(0, 1)), np.sqrt((se ** 2).sum(axis=(0, 1))) / 64
    expected = spectrum_to_rgb(flat(1.0 / (1 - rho)))
    print(f"pt furnace mode {mode}: {img} +- {img_se} expected {expected}", flush=True)
    assert np.all(np.abs(img - expected) < 4 * img_se + 2e-3 * expected)


@pytest.mark.parametrize("mode", [abi.SAMPLE_BOTH, abi.SAMPLE_BSDF, abi.SAMPLE_LIGHTS])
def test_gpu_pt_visible_emitter_every_pixel(gpu_lib, mode):
    L = 2.0
    w, h = 8, 6
    mean, se = gpu_batches(emitter_wall_scene(L), _params(w, h, 4, mode, abi.MIS_CONSISTENT), 8, 1024)
    expected = spectrum_to_rgb(flat(L))
    assert np.all(np.abs(mean - expected) < 5 * se + 2e-3 * expected), (mode, mean[0, :2], expected)


def test_gpu_pt_reference_pixel_truncation_quirk(gpu_lib):
    L = 2.0
    w, h = 8, 6
    mean, se = gpu_batches(emitter_wall_scene(L), _params(w, h, 4, abi.SAMPLE_LIGHTS, abi.MIS_REFERENCE), 8, 1024)
    expected = spectrum_to_rgb(flat(L))[None, None, :] * (1 + edge_pattern(w, h))[:, :, None]
    assert np.all(np.abs(mean - expected) < 5 * se + 2e-3 * expected)


def test_gpu_pt_rejects_media(gpu_lib):
    with pytest.raises(TTGpuError, match="surface scenes"):
        GpuRenderer(fog_furnace_scene(0.5), _params(4, 4, 8, abi.SAMPLE_BOTH, abi.MIS_CONSISTENT))
