"""Guiding building blocks on the CPU: the oracle pinned by closed forms and the
reference's own known-answer tests, and the library's host recorder checked
against the oracle.

* vMF mixture (src/distribution_mixture_models.cxx:88-150): normalisation of the
  density by cubature, sampler chi-square against the density.
* Incremental MAP-EM (:247-343): recovers a concentrated lobe, puts the mass of a two-lobe
  mixture on its modes, is invariant to weight scaling; the library's
  FP32 fit equals the oracle's FP64 fit on the same weighted records.
* kd-tree builder (src/path_guiding_tree.hxx:300-438): the reference's
  KdTreeBuilder* known answers (src/tests_guiding.cxx:221-320), and equal leaf
  partitions of library and oracle on random points.
"""
import numpy as np
import pytest

from naivetrace_amd import abi
from naivetrace_amd.renderer import build_kdtree, vmf_fit, vmf_init
import stats_util as S


def _mixture(mus, kappas, weights):
    m = abi.Vmf8()
    for k in range(8):
        mu = np.asarray(mus[k % len(mus)], float)
        mu = mu / np.linalg.norm(mu)
        for c in range(3):
            m.mu[k][c] = mu[c]
        m.kappa[k] = kappas[k % len(kappas)]
        m.w[k] = weights[k]
    return m


def _random_mixture(seed, kmax=40.0):
    rng = np.random.default_rng(seed)
    mus = rng.normal(size=(8, 3))
    w = rng.uniform(0.2, 1.0, 8)
    return _mixture(mus, rng.uniform(0.5, kmax, 8), w / w.sum())


def _sphere_integral(pdf, n=256):
    # Midpoint rule in (cos theta, phi): area element is exact, so a smooth density integrates accurately.
    z = (np.arange(n) + 0.5) / n * 2 - 1
    phi = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    Z, P = np.meshgrid(z, phi, indexing="ij")
    r = np.sqrt(1 - Z ** 2)
    d = np.stack([r * np.cos(P), r * np.sin(P), Z], axis=-1).reshape(-1, 3)
    return pdf(d).sum() * (2.0 / n) * (2 * np.pi / (2 * n))


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_vmf_normalised(oracle, seed):
    for m in (oracle.vmf_init(), _random_mixture(seed, kmax=12.0)):
        assert abs(_sphere_integral(lambda d: oracle.vmf_pdf(m, d), 512) - 1.0) < 2e-3


def test_oracle_vmf_sampler_chi_square(oracle):
    m = _random_mixture(3, kmax=25.0)
    rng = np.random.default_rng(5)
    n = 400000
    d = oracle.vmf_sample(m, rng.uniform(size=(n, 3)))
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    counts = np.bincount(S.cube_bin(d, 4), minlength=96).astype(float)
    probs = S.bin_probabilities(lambda x: oracle.vmf_pdf(m, x), 4, 32)
    pval, chi2, dof = S.chi_square_pvalue(counts, probs, n)
    assert pval > 1e-3, (chi2, dof)


def test_oracle_vmf_extreme_concentration_sampling(oracle):
    """Guiding.MovmfSampling (src/tests_guiding.cxx:639-676): samples stay finite and of unit length
    for the extreme concentrations the fit can clamp to."""
    m = _mixture([[0.83, -0.23, 0.5], [0.25, 0.97, 0.01]], [40.0, 1e-4], [0.999] + [1e-3 / 7] * 7)
    u = np.array([[0.119874, 0.395788, 8.51584e-08], [0.5, 0.5, 1 - 1e-9], [0.9999, 0.1, 1e-12]])
    d = oracle.vmf_sample(m, u)
    assert np.all(np.isfinite(d))
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)


def _svmf(oracle, mu, k, n, seed):
    return oracle.vmf_sample(_mixture([mu], [k], [1.0] + [0.0] * 7), np.random.default_rng(seed).uniform(size=(n, 3)))


def test_oracle_em_recovers_single_lobe(oracle):
    """One concentrated lobe of data: the heaviest fitted lobe sits on it with its concentration.
    (The MAP prior keeps pulling the minor lobes towards their initial directions, so only the
    dominant lobe is compared with the truth.)"""
    mu = np.array([0.3, 0.5, 0.8]) / np.linalg.norm([0.3, 0.5, 0.8])
    d = _svmf(oracle, mu, 20.0, 20000, 1)
    m = oracle.vmf_fit(oracle.vmf_init(), d, np.ones(len(d)), 20000)
    W, M, K = np.array(m.w), np.array(m.mu), np.array(m.kappa)
    i = W.argmax()
    assert W[i] > 0.85
    assert M[i] @ mu > np.cos(np.radians(4.0))
    assert abs(K[i] / 20.0 - 1) < 0.15


def test_oracle_em_two_lobes_mass_and_likelihood(oracle):
    truth = _mixture([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]], [25.0, 8.0], [0.7, 0.3] + [0.0] * 6)
    n = 60000
    d = oracle.vmf_sample(truth, np.random.default_rng(7).uniform(size=(n, 3)))
    m = oracle.vmf_fit(oracle.vmf_init(), d, np.ones(n), 10000)
    W, M = np.array(m.w), np.array(m.mu)
    for mu_t, w_t in (([0, 0, 1], 0.7), ([1, 0, 0], 0.3)):
        near = M @ np.asarray(mu_t, float) > 0.9
        assert abs(W[near].sum() - w_t) < 0.1, W
    x = d[:5000]
    kl_fit = np.mean(np.log(oracle.vmf_pdf(truth, x) / oracle.vmf_pdf(m, x)))
    kl_init = np.mean(np.log(oracle.vmf_pdf(truth, x) / oracle.vmf_pdf(oracle.vmf_init(), x)))
    assert kl_fit < 0.2 and kl_fit < 0.1 * kl_init, (kl_fit, kl_init)


def test_oracle_em_invariants(oracle):
    """Scaling all record weights leaves the fit unchanged (weights enter normalised by their
    mean); every fit is a valid mixture (unit means, clamped concentrations, weights summing to 1)."""
    rng = np.random.default_rng(3)
    d = _svmf(oracle, [0.0, 1.0, 0.0], 6.0, 3000, 2)
    w = rng.exponential(1.0, len(d))
    a = oracle.vmf_fit(oracle.vmf_init(), d, w, 3000)
    b = oracle.vmf_fit(oracle.vmf_init(), d, 37.5 * w, 3000)
    np.testing.assert_allclose(np.array(a.mu), np.array(b.mu), atol=1e-6)
    np.testing.assert_allclose(np.array(a.kappa), np.array(b.kappa), rtol=1e-5)
    np.testing.assert_allclose(np.array(a.w), np.array(b.w), atol=1e-6)
    for m in (a, b):
        np.testing.assert_allclose(np.linalg.norm(np.array(m.mu), axis=1), 1.0, atol=1e-6)
        assert np.all((np.array(m.kappa) >= 1e-4) & (np.array(m.kappa) <= 40.0))
        assert abs(np.sum(m.w) - 1.0) < 1e-6


def test_library_em_equals_oracle(oracle):
    """The recorder's FP32 EM (host code of the library) against the oracle's FP64 EM, same records."""
    truth = _mixture([[0.2, 0.3, 1.0], [-1.0, 0.4, 0.1], [0.0, -1.0, 0.2]], [15.0, 5.0, 30.0],
                     [0.5, 0.3, 0.2] + [0.0] * 5)
    rng = np.random.default_rng(11)
    n = 20000
    d = oracle.vmf_sample(truth, rng.uniform(size=(n, 3)))
    w = rng.exponential(1.0, n) * (rng.uniform(size=n) > 0.1)  # heavy spread and some zero weights
    d32 = d.astype(np.float32)
    ref = oracle.vmf_fit(oracle.vmf_init(), d32.astype(np.float64), w.astype(np.float32).astype(np.float64), 5000)
    lib = vmf_fit(vmf_init(), d32, w.astype(np.float32), 5000)
    np.testing.assert_allclose(np.array(lib.w), np.array(ref.w), atol=2e-4)
    np.testing.assert_allclose(np.array(lib.mu), np.array(ref.mu), atol=2e-4)
    np.testing.assert_allclose(np.array(lib.kappa), np.array(ref.kappa), rtol=2e-3, atol=1e-4)


def test_library_prior_equals_oracle(oracle):
    a, b = vmf_init(), oracle.vmf_init()
    np.testing.assert_array_equal(np.array(a.mu), np.array(b.mu))
    np.testing.assert_array_equal(np.array(a.kappa), 2.0)
    np.testing.assert_array_equal(np.array(a.w), 0.125)


CUBE = np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [1, 1, -1], [-1, -1, 1], [1, -1, 1], [-1, 1, 1], [1, 1, 1]], float)


@pytest.mark.parametrize("builder", ["oracle", "library"])
def test_kdtree_builder_known_answers(oracle, builder):
    """KdTreeBuilder, ..HandlesDegeneracy, ..MaxDepth, ..NumPoints (src/tests_guiding.cxx:221-320)."""
    build = oracle.kdtree_build if builder == "oracle" else build_kdtree
    leaf, n = build(CUBE, 1, 20)
    assert n == 8 and sorted(leaf.tolist()) == list(range(8))
    assert build(np.zeros((4, 3)), 1, 2)[1] == 2
    assert build(CUBE, 1, 2)[1] == 2
    assert build(CUBE, 4, 20)[1] == 2


def test_library_kdtree_equals_oracle(oracle):
    rng = np.random.default_rng(2)
    pts = rng.normal(size=(5000, 3)) * [3.0, 1.0, 0.2]
    for min_points, depth in ((100, 20), (7, 20), (1, 6)):
        la, na = build_kdtree(pts, min_points, depth)
        lb, nb = oracle.kdtree_build(pts, min_points, depth)
        assert na == nb
        # Leaf numbering is depth first in both; the partitions are identical.
        np.testing.assert_array_equal(la, lb)
        counts = np.bincount(la)
        assert counts.min() >= min(min_points, len(pts))
