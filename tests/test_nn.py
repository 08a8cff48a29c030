"""The learned guide's numpy oracle (oracle/ref_nn.py), pinned by central finite
differences of its own loss in float64: the mixture negative log-likelihood at
the raw outputs, and the whole network (hash grid, dense layers, biases)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import ref_nn  # noqa: E402


def _records(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.05, 0.95, (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi = rng.normal(size=(n, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    w = rng.exponential(1.0, n)
    return pos, nrm, wo, wi, w


def _random_params(net, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    L = net.L
    p = np.zeros(L.total)
    p[L.grid:L.w1] = rng.normal(0, 0.5, L.w1)
    for a, b, rows, cols, used in ((L.w1, L.w2, 64, 64, 38), (L.w2, L.w3, 64, 64, 64), (L.w3, L.w4, 64, 64, 64),
                                   (L.w4, L.b1, 96, 64, 64)):
        W = np.zeros((rows, cols))
        r_used = 80 if rows == 96 else rows
        W[:r_used, :used] = rng.normal(0, scale, (r_used, used))
        p[a:b] = W.reshape(-1)
    for a, b, m in ((L.b1, L.b2, 64), (L.b2, L.b3, 64), (L.b3, L.b4, 64), (L.b4, L.total, 80)):
        p[a:a + m] = rng.normal(0, 0.2, m)
    return p


def test_mixture_gradient_finite_differences():
    rng = np.random.default_rng(1)
    n = 6
    raw = rng.normal(0, 1.0, (n, 96))
    raw[:, 80:] = 0
    wi = rng.normal(size=(n, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    w = rng.exponential(1.0, n)
    _, g = ref_nn.Net.mixture_nll_grad(raw, wi, w)
    eps = 1e-6
    for i in range(n):
        for o in range(80):
            a, b = raw.copy(), raw.copy()
            a[i, o] += eps
            b[i, o] -= eps
            la, _ = ref_nn.Net.mixture_nll_grad(a, wi, w)
            lb, _ = ref_nn.Net.mixture_nll_grad(b, wi, w)
            fd = (la[i] - lb[i]) / (2 * eps)
            assert abs(fd - g[i, o]) < 1e-6 + 1e-5 * abs(fd), (i, o, fd, g[i, o])
    assert np.all(g[:, 80:] == 0)


def test_network_gradient_finite_differences():
    net = ref_nn.Net(levels=4, log2_table=8, base_res=3.0, max_res=12.0, bb_lo=(0, 0, 0), bb_hi=(1, 1, 1))
    p = _random_params(net, 2)
    pos, nrm, wo, wi, w = _records(12, 3)
    g, loss = net.gradient(p, pos, nrm, wo, wi, w, bf=False)
    assert np.isfinite(loss)
    L = net.L
    rng = np.random.default_rng(4)
    # Grid entries actually touched by the records, weights and biases of every layer (used rows/cols).
    touched = np.unique(np.concatenate([lv * L.T * 2 + 2 * net.corners(pos, lv)[0].reshape(-1) for lv in range(L.levels)]))
    picks = list(rng.choice(touched, 20, replace=False)) + list(rng.choice(touched + 1, 10, replace=False))
    for a, rows, cols, used_r, used_c in ((L.w1, 64, 64, 64, 38), (L.w2, 64, 64, 64, 64), (L.w3, 64, 64, 64, 64),
                                          (L.w4, 96, 64, 80, 64)):
        for _ in range(8):
            picks.append(a + rng.integers(used_r) * cols + rng.integers(used_c))
    for a, m in ((L.b1, 64), (L.b2, 64), (L.b3, 64), (L.b4, 80)):
        picks += list(a + rng.integers(0, m, 4))
    eps = 1e-6
    for i in picks:
        pa, pb = p.copy(), p.copy()
        pa[i] += eps
        pb[i] -= eps
        fd = (net.loss(pa, pos, nrm, wo, wi, w) - net.loss(pb, pos, nrm, wo, wi, w)) / (2 * eps)
        assert abs(fd - g[i]) < 1e-7 + 1e-4 * abs(fd), (i, fd, g[i])
    # Padding stays out: zero gradient for unused input columns and output rows.
    W1 = g[L.w1:L.w2].reshape(64, 64)
    W4 = g[L.w4:L.b1].reshape(96, 64)
    assert np.all(W1[:, 38:] == 0) and np.all(W4[80:] == 0) and np.all(g[L.b4 + 80:L.total] == 0)


def test_bf16_rounding():
    # bf16 keeps 7 fraction bits: spacing 2^-7 at 1.  Ties go to the even neighbour.
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9, 1 + 3 * 2 ** -9, 1.0078125, 3.0e38], np.float32)
    y = ref_nn.bf16(x)
    np.testing.assert_array_equal(y[:6], [1.0, 1.0, 1.015625, 1.0, 1.0078125, 1.0078125])
    assert np.all(np.isfinite(y))
