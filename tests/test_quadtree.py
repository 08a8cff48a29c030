"""Quadtree directional distribution on the CPU: the oracle pinned by the
reference's own known answers (src/tests_guiding.cxx:515-637, 679-691), closed
forms and brute force, and the library's host model checked against the oracle.

* sphere <-> square map (src/path_guiding.cxx:112-129): round trip, and the
  equal-area property (Archimedes: the Jacobian is 4 pi everywhere).
* builder (src/path_guiding_quadtree.cxx:32-88): QuadTreeBuilderTrivial,
  QuadTreeDivideIntoChildRanges and QuadTreeBuilderConsistency known answers; a
  single heavy point descends to depth 8 (29 nodes, density 4^7 there); every
  node's weight equals the brute-force sum over its region.
* adaptor (src/path_guiding_quadtree.hxx:209-315): the TreeAdaptor known answer
  (1 -> 5 -> 21 nodes), even splitting of a leaf to the threshold, collapse of
  light branches.
* hierarchical warping (src/path_guiding_quadtree.cxx:168-244): leaf frequencies
  and the position inside the leaf by chi-square against w_leaf / w_root and
  uniformity; the density integrates to 1.
* learned model (src/path_guiding.cxx:156-294): node statistics equal the
  two-pass mean / M2 of the pushed weights over each node's region; with counts
  <= 2 the sampling weights are the prior; round 1 has no exploration term and
  later rounds add more to rarely visited nodes; IterationUpdateAndBake carries
  counts scaled by the squared fractional size.
* library == oracle on random inputs for build, adapt, every model step and
  the direction filter's cone.
"""
import numpy as np
import pytest

from naivetrace_amd.renderer import QuadModel, quadtree_adapt, quadtree_build
import stats_util as S


def _unit(rng, n):
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def _regions(children):
    """(lo_u, lo_v, size, depth) of every node of a pre-order tree; root depth 1."""
    n = len(children)
    reg = np.zeros((n, 4))
    reg[0] = (0, 0, 1, 1)
    for i in range(n):
        if children[i][0] < 0:
            continue
        lo_u, lo_v, size, depth = reg[i]
        for c in range(4):
            k = children[i][c]
            reg[k] = (lo_u + (size / 2 if c & 1 else 0), lo_v + (size / 2 if c & 2 else 0), size / 2, depth + 1)
    return reg


def _leaves(children):
    return np.array([i for i in range(len(children)) if children[i][0] < 0])


# ---- map -------------------------------------------------------------------------------------

def test_oracle_map_round_trip_and_equal_area(oracle):
    import oracle_lib as O
    rng = np.random.default_rng(1)
    uv = rng.random((2000, 2)) * 0.998 + 0.001
    d = O.qt_square_to_sphere(uv)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    assert np.allclose(O.qt_sphere_to_square(d), uv, atol=1e-12)  # PolarMap (tests_guiding.cxx:679-691)
    # |d/du x d/dv| = 4 pi: the map preserves area up to a constant.
    h = 1e-6
    du = (O.qt_square_to_sphere(uv + [h, 0]) - O.qt_square_to_sphere(uv - [h, 0])) / (2 * h)
    dv = (O.qt_square_to_sphere(uv + [0, h]) - O.qt_square_to_sphere(uv - [0, h])) / (2 * h)
    jac = np.linalg.norm(np.cross(du, dv), axis=1)
    assert np.allclose(jac, 4 * np.pi, rtol=1e-5)
    # Child 0 is the bottom-left quadrant: u < 1/2 is the half-space y < 0, v < 1/2 is z < 0.
    p = O.qt_sphere_to_square(np.array([[0.0, -0.6, -0.8], [0.0, 0.6, 0.8]]))
    assert (p[0] < 0.5).all() and (p[1] > 0.5).all()


# ---- builder ---------------------------------------------------------------------------------

def test_oracle_builder_reference_known_answers(oracle):
    import oracle_lib as O
    # QuadTreeBuilderTrivial: no points -> a single leaf.
    ch, w = O.qt_build(np.zeros((0, 2)), np.zeros(0), 1.0)
    assert len(ch) == 1 and ch[0][0] == -1
    # QuadTreeDivideIntoChildRanges: 1, 2, 3, 4 points at the quadrant centres.
    centres = [(0.25, 0.25), (0.75, 0.25), (0.25, 0.75), (0.75, 0.75)]
    pts = [c for i, c in enumerate(centres) for _ in range(i + 1)]
    ch, w = O.qt_build(np.array(pts), np.ones(len(pts)), 0.45)
    assert ch.tolist()[0] == [1, 2, 3, 4] and (ch[1:] == -1).all()
    assert w.tolist() == [10, 1, 2, 3, 4]
    # QuadTreeBuilderConsistency: threshold just above one point -> leaves hold <= 1 point
    # unless at depth 8.
    rng = np.random.default_rng(3)
    pts = rng.random((100, 2))
    ch, w = O.qt_build(pts, np.ones(100), 1.00001 / 100)
    assert ch[0][0] >= 0
    reg = _regions(ch)
    for leaf in _leaves(ch):
        lo_u, lo_v, size, depth = reg[leaf]
        inside = ((pts[:, 0] >= lo_u) & (pts[:, 0] < lo_u + size) & (pts[:, 1] >= lo_v) & (pts[:, 1] < lo_v + size)).sum()
        assert inside == w[leaf]
        assert inside <= 1 or depth == 8


def test_oracle_builder_single_point_closed_form(oracle):
    import oracle_lib as O
    ch, w = O.qt_build(np.array([[0.3, 0.7]]), np.ones(1), 0.5)
    # Every node on the point's path splits (weight 1 > 0.5) until depth 8: 7 splits.
    assert len(ch) == 1 + 7 * 4
    reg = _regions(ch)
    assert reg[:, 3].max() == 8
    probs = np.where(w > 0, w, 0.0)
    pdf = O.qt_pdf(ch, probs, np.array([[0.3, 0.7], [0.9, 0.1]]))
    assert pdf[0] == pytest.approx(4.0 ** 7) and pdf[1] == 0.0


def test_oracle_builder_weights_are_region_sums(oracle):
    import oracle_lib as O
    rng = np.random.default_rng(4)
    pts = np.concatenate([rng.random((300, 2)), 0.05 * rng.random((300, 2)) + [0.6, 0.3]])
    wts = rng.random(600)
    ch, w = O.qt_build(pts, wts, 0.1)
    reg = _regions(ch)
    thr = 0.1 * wts.sum()
    for i in range(len(ch)):
        lo_u, lo_v, size, depth = reg[i]
        inside = (pts[:, 0] >= lo_u) & (pts[:, 0] < lo_u + size) & (pts[:, 1] >= lo_v) & (pts[:, 1] < lo_v + size)
        assert w[i] == pytest.approx(wts[inside].sum(), rel=1e-12, abs=1e-12)
        if ch[i][0] >= 0:
            assert w[i] > thr and depth < 8
        else:
            assert w[i] <= thr or depth == 8


# ---- adaptor ---------------------------------------------------------------------------------

def test_oracle_adaptor_reference_known_answers(oracle):
    import oracle_lib as O
    # TreeAdaptor (tests_guiding.cxx:610-637): one node of weight 16.
    ch, w = np.array([[-1, -1, -1, -1]]), np.array([16.0])
    ch, w, old, frac = O.qt_adapt(ch, w, 5 / 16)
    assert len(ch) == 5 and w.tolist() == [16, 4, 4, 4, 4] and frac.tolist() == [1, .5, .5, .5, .5]
    ch, w, old, frac = O.qt_adapt(ch, w, 2 / 16)
    assert len(ch) == 21
    # PushWeight of 10 at (0.75, 0.75) adds to the root.
    assert w[0] + 10 == 26


def test_oracle_adaptor_split_and_collapse(oracle):
    import oracle_lib as O
    # A single leaf of weight 1 at fraction 0.05: quarters 1/4, 1/16 keep splitting, 1/64 stops.
    ch, w, old, frac = O.qt_adapt(np.array([[-1, -1, -1, -1]]), np.array([1.0]), 0.05)
    leaves = _leaves(ch)
    assert len(ch) == 1 + 4 + 16 + 64 and len(leaves) == 64
    assert np.allclose(w[leaves], 1 / 64) and np.allclose(frac[leaves], 1 / 8) and (old == 0).all()
    # A branch at or below the threshold collapses into a leaf; heavy leaves split.
    ch = np.array([[1, 2, 3, 4], [5, 6, 7, 8], [-1] * 4, [-1] * 4, [-1] * 4, [-1] * 4, [-1] * 4, [-1] * 4, [-1] * 4])
    w = np.array([1.0, 0.04, 0.5, 0.3, 0.16, 0.01, 0.01, 0.01, 0.01])
    ch2, w2, old2, frac2 = O.qt_adapt(ch, w, 0.05)
    assert ch2[1].tolist() == [-1] * 4 and old2[1] == 1   # the 0.04 branch is now a leaf
    reg = _regions(ch2)
    # every leaf is below the threshold or at depth 8; the total is preserved
    for leaf in _leaves(ch2):
        assert w2[leaf] < 0.05 or reg[leaf][3] == 8 or (frac2[leaf] == 1 and w2[leaf] <= 0.05)
    assert np.isclose(w2[_leaves(ch2)].sum(), 1.0)


# ---- sampling --------------------------------------------------------------------------------

def _random_tree(O, seed, n=4000):
    rng = np.random.default_rng(seed)
    pts = np.concatenate([rng.random((n // 2, 2)), 0.1 * rng.random((n // 2, 2)) + [0.2, 0.7]])
    ch, w = O.qt_build(pts, rng.random(n), 0.02)
    probs = rng.random(len(ch)) + 0.05
    probs[rng.random(len(ch)) < 0.1] = 0.0  # some leaves never sampled
    return ch, O.qt_propagate(ch, probs)


def test_oracle_warping_chi_square(oracle):
    import oracle_lib as O
    ch, probs = _random_tree(O, 5)
    leaves = _leaves(ch)
    reg = _regions(ch)
    n = 200_000
    u = np.random.default_rng(6).random((n, 2))
    uv, pdf = O.qt_sample(ch, probs, u)
    # Which leaf each sample landed in, by brute force over the leaf regions.
    idx = np.full(n, -1)
    for k, leaf in enumerate(leaves):
        lo_u, lo_v, size, _ = reg[leaf]
        m = (uv[:, 0] >= lo_u) & (uv[:, 0] < lo_u + size) & (uv[:, 1] >= lo_v) & (uv[:, 1] < lo_v + size)
        idx[m] = k
    assert (idx >= 0).all()
    p_leaf = probs[leaves] / probs[0]
    counts = np.bincount(idx, minlength=len(leaves)).astype(float)
    assert counts[p_leaf == 0].sum() == 0
    pv, _, _ = S.chi_square_pvalue(counts, p_leaf, n)
    assert pv > 1e-4
    # Uniform inside the leaf: 4 x 4 sub-cells.
    rel = (uv - reg[leaves[idx]][:, :2]) / reg[leaves[idx]][:, 2:3]
    sub = (np.minimum((rel * 4).astype(int), 3) * [1, 4]).sum(1)
    pv2, _, _ = S.chi_square_pvalue(np.bincount(sub, minlength=16).astype(float), np.full(16, 1 / 16), n)
    assert pv2 > 1e-4
    # The returned density is the density at the point, and it integrates to 1.
    assert np.allclose(pdf, O.qt_pdf(ch, probs, uv), rtol=1e-12)
    sizes = reg[leaves][:, 2]
    assert np.sum(O.qt_pdf(ch, probs, reg[leaves][:, :2] + sizes[:, None] / 2) * sizes ** 2) == pytest.approx(1.0)


# ---- learned model ---------------------------------------------------------------------------

def _stats_brute(children, dirs, weights, O):
    uv = O.qt_sphere_to_square(dirs)
    reg = _regions(children)
    out = []
    for lo_u, lo_v, size, _ in reg:
        m = (uv[:, 0] >= lo_u) & (uv[:, 0] < lo_u + size) & (uv[:, 1] >= lo_v) & (uv[:, 1] < lo_v + size)
        x = weights[m]
        out.append((len(x), x.mean() if len(x) else 0.0, ((x - x.mean()) ** 2).sum() if len(x) else 0.0))
    return np.array(out)


def test_oracle_model_statistics_and_weights(oracle):
    import oracle_lib as O
    rng = np.random.default_rng(7)
    d0 = _unit(rng, 400)
    w0 = rng.random(400)
    m = O.QuadModel()
    m.initial_fit(d0, w0)
    learned = m.learned()
    st = _stats_brute(learned["children"], d0, w0, O)
    assert np.array_equal(learned["counts"], st[:, 0])
    assert np.allclose(learned["means"], st[:, 1]) and np.allclose(learned["m2"], st[:, 2])
    m.bake()
    s = m.sampled()
    # Bake: sampling weights are the node flux estimates (sum of weights / root count).
    assert np.allclose(s["probs"], st[:, 0] * st[:, 1] / 400)
    assert np.allclose(s["prior"], s["probs"])
    # IncrementalFit with only original samples pushes nothing: weights become prior x 1 (+ 0).
    m.incremental_fit(d0[:10], w0[:10], np.ones(10), 1)
    before = m.learned()
    assert np.array_equal(before["counts"], learned["counts"])
    # Filtered samples: counts grow; round 1 has no exploration term.
    d1, w1 = _unit(rng, 300), rng.random(300)
    m.incremental_fit(d1, w1, np.zeros(300), 1)
    l1 = m.learned()
    st1 = _stats_brute(l1["children"], np.concatenate([d0, d1]), np.concatenate([w0, w1]), O)
    assert np.array_equal(l1["counts"], st1[:, 0]) and np.allclose(l1["means"], st1[:, 1])
    assert np.allclose(l1["m2"], st1[:, 2])
    p1 = m.sampled()["probs"]
    leaves = _leaves(l1["children"])
    few = leaves[l1["counts"][leaves] <= 2]
    assert np.allclose(p1[few], s["prior"][few])  # mix factor 1: the prior alone
    # A later round adds sqrt(log(round) / (count + 1)) x root flux: more where counts are low.
    m2 = O.QuadModel()
    m2.initial_fit(d0, w0)
    m2.bake()
    m2.incremental_fit(d1, w1, np.zeros(300), 5)
    extra = m2.sampled()["probs"][leaves] - p1[leaves]
    flux = l1["means"][0]
    assert np.allclose(extra, flux * np.sqrt(np.log(5) / (l1["counts"][leaves] + 1)))


def test_oracle_update_and_bake_carries_scaled_counts(oracle):
    import oracle_lib as O
    rng = np.random.default_rng(8)
    d0, w0 = _unit(rng, 500), rng.random(500)
    m = O.QuadModel()
    m.initial_fit(d0, w0)
    m.bake()
    d1, w1 = _unit(rng, 500), rng.random(500) * 3
    m.incremental_fit(d1, w1, np.zeros(500), 1)
    before = m.learned()
    probs_before = m.sampled()["probs"]
    m.update_and_bake()
    s = m.sampled()
    after = m.learned()
    assert (after["counts"] == 0).all()                 # statistics reset
    assert np.array_equal(s["children"], after["children"])
    # The sampled tree is the adaptation of the flux estimates at fraction 0.05.
    flux = before["counts"] / before["counts"][0] * before["means"]
    ch, _, old, frac = O.qt_adapt(before["children"], flux, 0.05)
    assert np.array_equal(ch, s["children"])
    scale = frac ** 2
    n_new = np.floor(before["counts"][old] * scale)
    mix = 100 / (100 + np.maximum(n_new - 2, 0))
    leaves = _leaves(ch)
    heavy = leaves[n_new[leaves] >= 50]
    assert len(heavy) > 0
    means = scale * before["counts"][old] / before["counts"][0] * before["means"][old]
    sd = scale * before["counts"][old] / before["counts"][0] * np.sqrt(before["m2"][old] * scale) / (n_new - 1)
    want = scale * probs_before[old] * mix + (means + sd) * (1 - mix)
    assert np.allclose(s["probs"][heavy], want[heavy])
    assert np.allclose(s["prior"], s["probs"])


# ---- library vs oracle -----------------------------------------------------------------------

@pytest.mark.parametrize("seed", [11, 12])
def test_library_build_and_adapt_equal_oracle(oracle, seed):
    import oracle_lib as O
    rng = np.random.default_rng(seed)
    uv = np.concatenate([rng.random((3000, 2)), 0.03 * rng.random((1000, 2)) + [0.41, 0.83]]).astype(np.float32)
    w = rng.random(4000).astype(np.float32)
    ch, nw = quadtree_build(uv, w, 0.1)
    och, onw = O.qt_build(uv.astype(np.float64), w.astype(np.float64), 0.1)
    assert np.array_equal(ch, och)
    assert np.allclose(nw, onw, rtol=1e-6)
    ach, aw, aold, afrac = quadtree_adapt(ch, nw, 0.05)
    boch, bow, bold, bfrac = O.qt_adapt(och, onw, 0.05)
    assert np.array_equal(ach, boch) and np.array_equal(aold, bold)
    assert np.allclose(aw, bow, rtol=1e-6) and np.array_equal(afrac, bfrac)
    # reference known answers through the library as well
    ch, w2 = quadtree_build(np.zeros((0, 2)), np.zeros(0), 1.0)
    assert len(ch) == 1
    ch, w2, _, _ = quadtree_adapt(np.array([[-1] * 4]), np.array([16.0]), 5 / 16)
    ch, w2, _, _ = quadtree_adapt(ch, w2, 2 / 16)
    assert len(ch) == 21


def test_library_model_equals_oracle(oracle):
    import oracle_lib as O
    rng = np.random.default_rng(13)
    lib, ref = QuadModel(), O.QuadModel()

    def both(fn, *args):
        getattr(lib, fn)(*args)
        getattr(ref, fn)(*args)

    def compare():
        a, b = lib.learned(), ref.learned()
        assert np.array_equal(a["children"], b["children"])
        assert np.array_equal(a["counts"], b["counts"].astype(np.int64))
        assert np.allclose(a["means"], b["means"], rtol=1e-9, atol=1e-12)
        assert np.allclose(a["m2"], b["m2"], rtol=1e-7, atol=1e-10)
        a, b = lib.sampled(), ref.sampled()
        assert np.array_equal(a["children"], b["children"])
        for k in ("means", "stddev"):
            assert np.allclose(a[k], b[k], rtol=1e-6), k
        for k in ("probs", "prior"):
            assert np.allclose(a[k], b[k], rtol=2e-5, atol=1e-30), k

    # directions concentrated around two lobes, as a cell's incident radiance would be
    def draw(n):
        d = np.concatenate([_unit(rng, n // 2), _unit(rng, n - n // 2) * 0.15 + [0.0, 0.6, 0.8]])
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        w = np.where(d[:, 2] > 0.5, 5.0, 0.2) * rng.random(n)
        return d.astype(np.float32).astype(np.float64), w.astype(np.float32).astype(np.float64)

    d, w = draw(2000)
    both("initial_fit", d, w)
    both("bake")
    compare()
    for sweep in range(3):
        for _ in range(2 ** (sweep + 1)):
            d, w = draw(3000)
            orig = rng.random(3000) < 0.5
            both("incremental_fit", d, w, orig, 1)
            compare()
        both("update_and_bake")
        compare()
    both("incremental_fit", *draw(1000), np.zeros(1000), 4)
    compare()
    # The direction filter: a uniform direction in the cone of the oracle's half-angle.
    for _ in range(5):
        dd = _unit(rng, 1)[0].astype(np.float32)
        ang = ref.filter_angle(dd.astype(np.float64))
        outs = np.array([lib.filter_direction(dd, rng.random(), rng.random()) for _ in range(400)])
        cosines = outs @ dd.astype(np.float64)
        assert np.allclose(np.linalg.norm(outs, axis=1), 1, atol=1e-5)
        assert cosines.min() >= np.cos(ang) - 1e-5
        # uniform in the cap: E[cos] = (1 + cos a) / 2
        assert abs(cosines.mean() - (1 + np.cos(ang)) / 2) < 5 * (1 - np.cos(ang)) / np.sqrt(12 * 400) + 1e-6


def test_library_quadtree_rejects_malformed_trees(gpu_free_lib=None):
    from naivetrace_amd import renderer as R
    with pytest.raises(R.TTGpuError):
        quadtree_adapt(np.array([[1, 2, 3, -1], [-1] * 4, [-1] * 4, [-1] * 4]), np.ones(4), 0.1)
    with pytest.raises(R.TTGpuError):
        quadtree_adapt(np.array([[0, 1, 2, 3]] + [[-1] * 4] * 3), np.ones(4), 0.1)
    with pytest.raises(R.TTGpuError):
        quadtree_build(np.zeros((2, 2)), np.ones(2), 0.1, capacity=0)
    m = QuadModel()
    with pytest.raises(R.TTGpuError):
        m.incremental_fit(np.array([[0, 0, 1.0]]), np.ones(1), np.zeros(1), 0)
