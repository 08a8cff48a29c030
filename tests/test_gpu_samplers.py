"""GPU BSDF and phase-function samplers against the oracle.

* Distribution: GPU samples binned on a cube map are chi-square tested against
  bin probabilities integrated from the oracle's FP64 pdf (the reference's
  SamplingConsistencyTest methodology, src/tests_sampling.cxx:540-770).
* Point-wise: GPU Evaluate (value and pdf) equals the oracle's to FP32
  tolerance on random direction pairs.
* Delta lobes: reflection/refraction probabilities and weights.
"""
import numpy as np
import pytest

from naivetrace_amd import abi
from naivetrace_amd.renderer import sample_shader, eval_shader, sample_phase
import stats_util as S
from test_oracle import SHADER_CASES, PHASE_CASES, _shader, _wi, BINS, NG

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,kw,theta,rough", SHADER_CASES, ids=[c[0] for c in SHADER_CASES])
def test_gpu_shader_distribution(gpu_lib, oracle, name, kw, theta, rough):
    sh = _shader(**kw)
    wi = _wi(theta)
    n = 1000000
    d, v, p = sample_shader(sh, wi, NG, NG, BINS, rough, 31, n)
    assert np.all(np.isfinite(d)) and np.all(np.isfinite(v)) and np.all(np.isfinite(p))
    nd = p >= 0
    counts = np.bincount(S.cube_bin(d[nd], 4), minlength=96).astype(np.float64)
    probs = S.bin_probabilities(lambda dirs: oracle.eval_shader(sh, wi, NG, NG, BINS, rough, dirs)[1], 4, 32)
    pval, chi2, dof = S.chi_square_pvalue(counts, probs, n)
    print(f"{name}: chi2 {chi2:.1f} dof {dof} p {pval:.3g}", flush=True)
    assert pval > 1e-3
    assert S.bins_within_sigma(counts, probs, n, 5.0, prob_err=0.002).mean() > 0.97


@pytest.mark.parametrize("name,kw,theta,rough", SHADER_CASES, ids=[c[0] for c in SHADER_CASES])
def test_gpu_shader_evaluate_matches_oracle(gpu_lib, oracle, name, kw, theta, rough):
    sh = _shader(**kw)
    wi = _wi(theta)
    rng = np.random.default_rng(7)
    dirs = rng.normal(size=(4000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    v, p = eval_shader(sh, wi, NG, NG, BINS, rough, dirs)
    vr, pr = oracle.eval_shader(sh, wi, NG, NG, BINS, rough, dirs.astype(np.float32).astype(np.float64))
    # FP32 against FP64: relative 1e-3 plus an absolute floor near zero / at lobe boundaries.
    ok_v = np.abs(v - vr) <= 1e-3 * np.abs(vr) + 1e-5
    ok_p = np.abs(p - pr) <= 1e-3 * np.abs(pr) + 1e-5
    assert ok_v.mean() > 0.995 and ok_p.mean() > 0.995, (ok_v.mean(), ok_p.mean())


def test_gpu_specular_deltas(gpu_lib):
    mirror = _shader(abi.SHADER_MIRROR, reflectance=0.9)
    wi = _wi(40.)
    d, v, p = sample_shader(mirror, wi, NG, NG, BINS, 0.0, 1, 1000)
    assert np.all(p < 0)
    np.testing.assert_allclose(d, np.tile([-wi[0], -wi[1], wi[2]], (1000, 1)), atol=1e-6)
    glass = _shader(abi.SHADER_DIELECTRIC, ior=1.5)
    d, v, p = sample_shader(glass, wi, NG, NG, BINS, 0.0, 2, 1000000)
    assert np.all(p < 0)
    assert abs((d[:, 2] > 0).mean() - 0.1) < 0.002
    w = v[:, 0] * np.abs(d[:, 2]) / np.abs(p)
    R = 0.048306
    assert abs(w.mean() - (R + (1 - R) * (1 / 1.5) ** 2)) < 0.003


@pytest.mark.parametrize("kind,g", PHASE_CASES)
def test_gpu_phase_distribution(gpu_lib, oracle, kind, g):
    wi = _wi(50.)
    n = 1000000
    d, v, p = sample_phase(kind, g, wi, 9, n)
    counts = np.bincount(S.cube_bin(d, 4), minlength=96).astype(np.float64)
    probs = S.bin_probabilities(lambda dirs: np.array([oracle.eval_phase(kind, g, wi, x) for x in dirs]), 4, 24)
    pval, chi2, dof = S.chi_square_pvalue(counts, probs, n)
    assert pval > 1e-3, (chi2, dof)
    vr = np.array([oracle.eval_phase(kind, g, wi, x) for x in d[:2000].astype(np.float64)])
    np.testing.assert_allclose(v[:2000], vr, rtol=2e-3, atol=1e-6)
