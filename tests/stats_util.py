"""Statistical harness for sampler tests, restated from the reference's
SamplingConsistencyTest (src/tests_sampling.cxx:540-770) and
tests_stats.cxx:43-208: directions are binned on a cube map (6 faces x n x n
cells), expected bin probabilities are integrated numerically from a pdf
function, and the counts are judged with a chi-square test plus a per-bin
binomial sigma test.  Delta peaks are counted separately."""
import numpy as np
from scipy import stats


def cube_bin(dirs, n):
    d = np.asarray(dirs, dtype=np.float64)
    a = np.abs(d)
    face_axis = np.argmax(a, axis=1)
    sign = np.take_along_axis(d, face_axis[:, None], 1)[:, 0] >= 0
    face = face_axis * 2 + (~sign).astype(int)
    m = np.take_along_axis(a, face_axis[:, None], 1)[:, 0]
    other = np.array([[1, 2], [2, 0], [0, 1]])[face_axis]
    u = np.take_along_axis(d, other[:, :1], 1)[:, 0] / m
    v = np.take_along_axis(d, other[:, 1:], 1)[:, 0] / m
    iu = np.clip(((u + 1) * 0.5 * n).astype(int), 0, n - 1)
    iv = np.clip(((v + 1) * 0.5 * n).astype(int), 0, n - 1)
    return (face * n + iu) * n + iv


def cell_directions(n, sub):
    """Midpoints of `sub`x`sub` sub-cells of every cube-map cell and their solid angles.
    Returns dirs (6*n*n, sub*sub, 3) and weights (6*n*n, sub*sub)."""
    k = n * sub
    g = (np.arange(k) + 0.5) / k * 2 - 1
    U, V = np.meshgrid(g, g, indexing="ij")
    du = 2.0 / k
    jac = du * du / (1 + U ** 2 + V ** 2) ** 1.5
    dirs = np.zeros((6, k, k, 3))
    for face in range(6):
        ax = face // 2
        s = 1.0 if face % 2 == 0 else -1.0
        o = [[1, 2], [2, 0], [0, 1]][ax]
        dirs[face, :, :, ax] = s
        dirs[face, :, :, o[0]] = U
        dirs[face, :, :, o[1]] = V
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    dirs = dirs.reshape(6, n, sub, n, sub, 3).transpose(0, 1, 3, 2, 4, 5).reshape(6 * n * n, sub * sub, 3)
    w = np.broadcast_to(jac[None], (6, k, k)).reshape(6, n, sub, n, sub).transpose(0, 1, 3, 2, 4).reshape(6 * n * n, sub * sub)
    return dirs, w


def bin_probabilities(pdf_fn, n=4, sub=24):
    dirs, w = cell_directions(n, sub)
    p = pdf_fn(dirs.reshape(-1, 3)).reshape(w.shape)
    return (p * w).sum(axis=1)


def chi_square_pvalue(counts, probs, total, min_expected=5.0):
    exp = probs * total
    keep = exp >= min_expected
    obs_k, exp_k = counts[keep], exp[keep]
    obs_rest = counts[~keep].sum() + max(0.0, total - counts.sum() - 0)
    exp_rest = exp[~keep].sum()
    obs = list(obs_k)
    ex = list(exp_k)
    if exp_rest >= min_expected:
        obs.append(counts[~keep].sum())
        ex.append(exp_rest)
    obs = np.array(obs, dtype=np.float64)
    ex = np.array(ex, dtype=np.float64)
    ex *= obs.sum() / ex.sum()
    chi2 = ((obs - ex) ** 2 / ex).sum()
    dof = max(1, len(obs) - 1)
    return float(stats.chi2.sf(chi2, dof)), chi2, dof


def bins_within_sigma(counts, probs, total, nsigma=5.0, prob_err=None):
    """Per-bin binomial test (CheckNumberOfSamplesInBin, tests_sampling.cxx:586-604)."""
    exp = probs * total
    sigma = np.sqrt(np.maximum(total * probs * (1 - probs), 1e-12))
    slack = 0 if prob_err is None else prob_err * total
    return np.abs(counts - exp) <= nsigma * sigma + slack + 1.0
