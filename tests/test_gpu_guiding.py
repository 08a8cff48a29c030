"""Guided integrator (ptg) on the GPU, through the C ABI.

* The GPU's vMF mixture sampler: chi-square of its samples against the oracle's FP64
  density, and its density against the oracle's point-wise.
* Unbiasedness: the guided estimator (BSDF/vMF one-sample mixture, power-heuristic MIS)
  converges to the same image as pt2 with selection-consistent MIS: per-pixel 3-sigma
  parity with the oracle's converged Cornell render (src/renderingalgorithms_pathtracing_guided.cxx:473-544).
* Training schedule (:1110-1160): one initial pass plus sweeps of 2, 4, .. 1-spp passes; the
  kd-tree adapts (more than one cell) and records flow every pass.
* Guide quality: on the Cornell box with a hidden light, the guided render has lower
  relMSE than pt2 at equal samples per pixel.
"""
import os

import numpy as np
import pytest

from naivetrace_amd import abi, configs
from naivetrace_amd.renderer import GpuRenderer, vmf_pdf_gpu, vmf_sample_gpu
import stats_util as S
from test_gpu_parity import GOLDEN, _zscores
from test_guiding import _random_mixture

pytestmark = pytest.mark.gpu


def test_gpu_vmf_sampler_matches_oracle(gpu_lib, oracle):
    for seed in (3, 4):
        m = _random_mixture(seed, kmax=30.0)
        n = 1000000
        d, p = vmf_sample_gpu(m, 17, n)
        assert np.all(np.isfinite(d)) and np.all(p > 0)
        np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
        counts = np.bincount(S.cube_bin(d, 4), minlength=96).astype(float)
        probs = S.bin_probabilities(lambda x: oracle.vmf_pdf(m, x), 4, 32)
        pval, chi2, dof = S.chi_square_pvalue(counts, probs, n)
        print(f"vmf seed {seed}: chi2 {chi2:.1f} dof {dof} p {pval:.3g}", flush=True)
        assert pval > 1e-3
        x = d[:20000].astype(np.float64)
        np.testing.assert_allclose(vmf_pdf_gpu(m, x), oracle.vmf_pdf(m, x), rtol=2e-4, atol=1e-6)


def _ptg_params(w, h, depth, guide_max_spp, spp, seed=23):
    p = abi.default_params(w, h, spp=spp, max_depth=depth, seed=seed, algo=abi.ALGO_PTG)
    p.guide_kind = abi.GUIDE_VMF_TREE
    p.guide_max_spp = guide_max_spp
    p.guide_subdiv_factor = 100
    return p


def test_gpu_ptg_schedule_and_recorder(gpu_lib):
    s = configs.cornell_hidden_light(64, 64)
    p = _ptg_params(64, 64, 8, 8, 16)
    calls = []
    with GpuRenderer(s, p) as r:
        r.set_pass_callback(lambda c: calls.append(c))
        assert r.run() == 0
        info = r.guide_info()
        _, _, n = r.framebuffer()
    print("guide info", info, flush=True)
    # 1 initial pass + sweeps of 2, 4, 8 passes; then the render schedule 1 + 2 + 4 + 8 + 1 = 16 spp.
    assert info["rounds"] == 1 + 2 + 4 + 8
    assert info["cells"] > 1 and info["records_last_pass"] > 0 and info["records_dropped"] == 0
    assert np.all(n == 16)
    assert len(calls) == 1 + 3 + 5


def test_gpu_ptg_unbiased_cornell(gpu_lib):
    f = np.load(os.path.join(GOLDEN, "cornell32_consistent_oracle.npz"))
    w, h, depth, spp = int(f["width"]), int(f["height"]), int(f["max_depth"]), int(f["spp"])
    s = configs.cornell_box(w, h)
    p = _ptg_params(w, h, depth, 16, spp)
    with GpuRenderer(s, p) as r:
        r.run()
        rgb, sq, n = r.framebuffer()
    assert np.all(n == spp)
    mean = rgb / spp
    var = np.maximum(sq / spp - mean ** 2, 0) * spp / (spp - 1)
    z = _zscores(mean, np.sqrt(var / spp), f["mean"], f["stderr"])
    frac3 = float(np.mean(np.abs(z) > 3))
    npix = w * h
    img, img_ref = mean.mean(axis=(0, 1)), f["mean"].mean(axis=(0, 1))
    img_se = np.sqrt((var / spp).sum(axis=(0, 1)) + (f["stderr"] ** 2).sum(axis=(0, 1))) / npix
    print(f"ptg cornell: mean z {z.mean():+.4f} |z|>3 {frac3:.4f} image {img} oracle {img_ref} +- {img_se}", flush=True)
    assert frac3 < 0.015
    assert abs(z.mean()) < 5.0 / np.sqrt(z.size) + 0.05
    assert np.all(np.abs(img - img_ref) < 4 * img_se + 1e-3 * np.abs(img_ref))


def _relmse(img, ref):
    lum = lambda x: x @ np.array([0.2126, 0.7152, 0.0722])
    a, b = lum(img), lum(ref)
    return float(np.mean((a - b) ** 2 / (b ** 2 + 1e-2)))


def test_gpu_ptg_reduces_error_hidden_light(gpu_lib):
    w = h = 64
    s = configs.cornell_hidden_light(w, h)
    pref = abi.default_params(w, h, max_depth=8, seed=99)
    pref.mis_mode = abi.MIS_CONSISTENT
    with GpuRenderer(s, pref) as r:
        r.render_spp(32768)
        ref = r.mean_image()
    spp = 64
    pp = abi.default_params(w, h, spp=spp, max_depth=8, seed=5)
    pp.mis_mode = abi.MIS_CONSISTENT
    with GpuRenderer(s, pp) as r:
        r.run()
        plain = r.mean_image()
    with GpuRenderer(s, _ptg_params(w, h, 8, 64, spp, seed=6)) as r:
        r.run()
        guided = r.mean_image()
    e_plain, e_guided = _relmse(plain, ref), _relmse(guided, ref)
    print(f"hidden light {spp} spp: relMSE plain {e_plain:.5f} guided {e_guided:.5f} ratio {e_plain / e_guided:.2f}", flush=True)
    assert e_guided < e_plain
