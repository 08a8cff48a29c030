// Quadtree directional distribution (host side).  See quadtree.hpp for the
// reference files followed.
#include "quadtree.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <functional>
#include <stdexcept>

namespace ttgpu {
namespace guide {
namespace quad {

namespace {
constexpr double kPi = 3.14159265358979323846;
const float kLessThan1 = std::nextafter(1.f, 0.f);

int ChildNum(Point p, const float lo[2], float size) {
  const float cx = lo[0] + 0.5f * size, cy = lo[1] + 0.5f * size;
  return (p.u < cx ? 0 : 1) | (p.v < cy ? 0 : 2);
}

int NewNode(Tree &t) {
  if ((int)t.child.size() >= kMaxNodes) throw std::runtime_error("quadtree exceeds 65534 nodes");
  t.child.push_back({-1, -1, -1, -1});
  return (int)t.child.size() - 1;
}

void PushWeight(const Tree &t, Stats &s, Point p, float w) {
  Descent d(p, t);
  for (;;) {
    s.Add(d.node, w);
    if (d.Leaf(t)) break;
    d.Step(t);
  }
}
}  // namespace

void Descent::Step(const Tree &t) {
  const int c = ChildNum(p, lo, size);
  size *= 0.5f;
  if (c & 1) lo[0] += size;
  if (c & 2) lo[1] += size;
  node = t.child[node][c];
  ++depth;
}

Point SphereToTree(const float d[3]) {
  const double v = ((double)d[2] + 1.) * 0.5;
  const double u = (std::atan2((double)d[1], (double)d[0]) / kPi + 1.) * 0.5;
  return {(float)u, (float)v};
}

void TreeToSphere(Point p, float out[3]) {
  const double z = (double)p.v * 2. - 1.;
  const double rho = std::sqrt(std::max(0., 1. - z * z));
  const double phi = ((double)p.u * 2. - 1.) * kPi;
  out[0] = (float)(rho * std::cos(phi));
  out[1] = (float)(rho * std::sin(phi));
  out[2] = (float)z;
}

Tree Build(const std::vector<Point> &pts, const std::vector<float> &w, float fraction, std::vector<float> &node_w) {
  if (pts.size() != w.size()) throw std::invalid_argument("quadtree builder: points and weights differ in length");
  Tree t;
  t.child.clear();
  node_w.clear();
  double total = 0.;
  for (float x : w) total += x;
  const float threshold = fraction * (float)total;
  std::vector<int> idx(pts.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
  std::function<int(int, int, const float *, float, int)> rec = [&](int b, int e, const float *lo, float size, int depth) {
    double sum = 0.;
    for (int k = b; k < e; ++k) sum += w[idx[k]];
    const float ws = (float)sum;
    const int node = NewNode(t);
    node_w.push_back(ws);
    if (ws > threshold && depth < kMaxDepth) {
      // Stable grouping by child number, then one recursion per child in order 0..3.
      std::vector<int> tmp(idx.begin() + b, idx.begin() + e);
      int start[5] = {0, 0, 0, 0, 0};
      for (int i : tmp) ++start[ChildNum(pts[i], lo, size) + 1];
      for (int c = 0; c < 4; ++c) start[c + 1] += start[c];
      int fill[4] = {start[0], start[1], start[2], start[3]};
      for (int i : tmp) idx[b + fill[ChildNum(pts[i], lo, size)]++] = i;
      const float half = 0.5f * size;
      for (int c = 0; c < 4; ++c) {
        const float clo[2] = {lo[0] + ((c & 1) ? half : 0.f), lo[1] + ((c & 2) ? half : 0.f)};
        const int ch = rec(b + start[c], b + start[c + 1], clo, half, depth + 1);
        t.child[node][c] = ch;
      }
    }
    return node;
  };
  const float lo0[2] = {0.f, 0.f};
  t.root = rec(0, (int)pts.size(), lo0, 1.f, 1);
  return t;
}

Tree Adapt(const Tree &t, const std::vector<float> &w, float fraction, std::vector<float> &new_w,
           std::vector<NewToOld> &mapping) {
  if ((int)w.size() != t.NumNodes()) throw std::invalid_argument("quadtree adaptor: one weight per node expected");
  Tree nt;
  nt.child.clear();
  new_w.clear();
  mapping.clear();
  const float threshold = fraction * w[t.root];
  auto alloc = [&](float weight, NewToOld old) {
    const int n = NewNode(nt);
    new_w.push_back(weight);
    mapping.push_back(old);
    return n;
  };
  // A leaf above the threshold is split evenly until the quarters fall below it.
  std::function<int(float, int, NewToOld)> split_leaf = [&](float weight, int depth, NewToOld old) {
    if (weight < threshold || depth >= kMaxDepth) return alloc(weight, old);
    const int n = alloc(weight, old);
    for (int c = 0; c < 4; ++c) {
      const int ch = split_leaf(weight * 0.25f, depth + 1, {old.idx, 0.5f * old.fractional_size});
      nt.child[n][c] = ch;
    }
    return n;
  };
  std::function<int(int, int)> adapt = [&](int node, int depth) {
    const float weight = w[node];
    if (depth < kMaxDepth && weight > threshold) {
      if (t.IsLeaf(node)) return split_leaf(weight, depth, {node, 1.f});
      const int n = alloc(weight, {node, 1.f});
      for (int c = 0; c < 4; ++c) {
        const int ch = adapt(t.child[node][c], depth + 1);
        nt.child[n][c] = ch;
      }
      return n;
    }
    return alloc(weight, {node, 1.f});  // a branch at or below the threshold collapses into a leaf
  };
  nt.root = adapt(t.root, 1);
  return nt;
}

void Propagate(const Tree &t, std::vector<float> &w) {
  std::function<void(int)> rec = [&](int n) {
    if (t.IsLeaf(n)) return;
    float s = 0.f;
    for (int c = 0; c < 4; ++c) {
      rec(t.child[n][c]);
      s += w[t.child[n][c]];
    }
    w[n] = s;
  };
  rec(t.root);
}

std::pair<Point, float> Sample(const Tree &t, const std::vector<float> &w, float r0, float r1) {
  r0 = std::min(r0, kLessThan1);
  r1 = std::min(r1, kLessThan1);
  int node = t.root;
  float lo[2] = {0.f, 0.f}, size = 1.f;
  while (!t.IsLeaf(node)) {
    const auto &ch = t.child[node];
    const float cw[4] = {w[ch[0]], w[ch[1]], w[ch[2]], w[ch[3]]};
    // left / right from the marginal, then bottom / top conditional on the column; each random
    // number is rescaled to [0, 1) within the chosen interval.
    const float wl = cw[0] + cw[2], wr = cw[1] + cw[3];
    const float s0 = r0 * (wl + wr);
    int c = s0 < wl ? 0 : 1;
    r0 = s0 < wl ? s0 / (wl + FLT_EPSILON) : (s0 - wl) / (wr + FLT_EPSILON);
    const float wb = c ? cw[1] : cw[0], wt = c ? cw[3] : cw[2];
    const float s1 = r1 * (wb + wt);
    c |= s1 < wb ? 0 : 2;
    r1 = s1 < wb ? s1 / (wb + FLT_EPSILON) : (s1 - wb) / (wt + FLT_EPSILON);
    r0 = std::min(r0, kLessThan1);
    r1 = std::min(r1, kLessThan1);
    size *= 0.5f;
    if (c & 1) lo[0] += size;
    if (c & 2) lo[1] += size;
    node = ch[c];
  }
  const float pdf = w[node] / (w[t.root] * (size * size));
  Point p{size * r0 + lo[0], size * r1 + lo[1]};
  p.u = std::max(std::min(p.u, (lo[0] + size) * kLessThan1), lo[0]);
  p.v = std::max(std::min(p.v, (lo[1] + size) * kLessThan1), lo[1]);
  return {p, pdf};
}

float Pdf(const Tree &t, const std::vector<float> &w, Point p) {
  Descent d(p, t);
  while (!d.Leaf(t)) d.Step(t);
  return w[d.node] / (w[t.root] * (d.size * d.size));
}

void Stats::Add(int i, double x) {
  ++n[i];
  const double delta = x - mean[i];
  mean[i] += delta / (double)n[i];
  m2[i] += delta * (x - mean[i]);
}

double Stats::MeanErr(int i, double fill, int64_t min_count) const {
  return n[i] >= min_count ? std::sqrt(m2[i]) / (double)(n[i] - 1) : fill;
}

double Sampled::Pdf(const float dir[3]) const {
  return (double)quad::Pdf(tree, probs, SphereToTree(dir)) * kJacobianInv;
}

int Sampled::LeafDepth(const float dir[3]) const {
  Descent d(SphereToTree(dir), tree);
  while (!d.Leaf(tree)) d.Step(tree);
  return d.depth;
}

void Sampled::FilterDirection(const float dir[3], double u0, double u1, float out[3]) const {
  float footprint = 0.5f;
  for (int k = LeafDepth(dir); k > 0; --k) footprint *= 0.5f;
  footprint *= 3.5449077018110318f;  // sqrt(4 pi)
  const double cos_open = (double)std::cos(footprint);
  const double z = 1. - u1 * (1. - cos_open);
  const double s = std::sqrt(std::max(0., 1. - z * z));
  const double om = 2. * kPi * u0;
  const float off[3] = {(float)(std::cos(om) * s), (float)(std::sin(om) * s), (float)z};
  // Duff et al. 2017 frame with dir as z.
  const float sign = std::copysign(1.f, dir[2]);
  const float a = -1.f / (sign + dir[2]);
  const float b = dir[0] * dir[1] * a;
  const float fx[3] = {1.f + sign * dir[0] * dir[0] * a, sign * b, -sign * dir[0]};
  const float fy[3] = {b, sign + dir[1] * dir[1] * a, -dir[1]};
  for (int c = 0; c < 3; ++c) out[c] = fx[c] * off[0] + fy[c] * off[1] + dir[c] * off[2];
}

void Learned::InitialFit(const std::vector<Sample1> &s) {
  std::vector<Point> pts(s.size());
  std::vector<float> w(s.size()), node_w;
  for (size_t i = 0; i < s.size(); ++i) {
    pts[i] = SphereToTree(s[i].dir);
    w[i] = s[i].weight;
  }
  tree = Build(pts, w, 0.1f, node_w);
  stats = Stats(tree.NumNodes());
  for (size_t i = 0; i < s.size(); ++i) PushWeight(tree, stats, pts[i], w[i]);
}

namespace {
double MixFactor(int64_t n) {  // good_sample_count = 100, min_sample_count = 2
  return 100. / (100. + (double)std::max<int64_t>(n - 2, 0));
}
}  // namespace

void Learned::IncrementalFit(const std::vector<Sample1> &s, int round, Sampled &dst) {
  if (round < 1) throw std::invalid_argument("quadtree incremental fit: round must be >= 1");
  for (const Sample1 &x : s)
    if (!x.original) PushWeight(tree, stats, SphereToTree(x.dir), x.weight);
  const int nn = tree.NumNodes();
  if ((int)dst.prior.size() != nn) throw std::logic_error("quadtree incremental fit: sampled and learned trees differ");
  const int64_t root_count = stats.n[tree.root];
  if (root_count <= 0) return;
  std::vector<double> rc(nn), means(nn), stddev(nn);
  for (int i = 0; i < nn; ++i) {
    rc[i] = (double)stats.n[i] / (double)root_count;
    means[i] = rc[i] * stats.mean[i];
    stddev[i] = rc[i] * stats.MeanErr(i, kLargeFloat, 2);
  }
  const double total_flux = means[tree.root];
  const double log_round = std::log((double)round);
  dst.probs.resize(nn);
  for (int i = 0; i < nn; ++i) {
    const double mix = MixFactor(stats.n[i]);
    dst.probs[i] = (float)((double)dst.prior[i] * mix + (means[i] + stddev[i]) * (1. - mix));
    // Exploration term after the upper-confidence-bound bandit rule.
    dst.probs[i] += (float)total_flux * (float)std::sqrt(log_round / (double)(stats.n[i] + 1));
  }
  Propagate(tree, dst.probs);
}

Sampled Learned::IterationUpdateAndBake(const Sampled &prev) {
  if (!(stats.n[tree.root] > 10)) return prev;
  const int on = tree.NumNodes();
  const double root_count = (double)stats.n[tree.root];
  std::vector<double> old_rc(on);
  std::vector<float> old_flux(on);
  for (int i = 0; i < on; ++i) {
    old_rc[i] = (double)stats.n[i] / root_count;
    old_flux[i] = (float)(old_rc[i] * stats.mean[i]);
  }
  std::vector<float> new_w;
  std::vector<NewToOld> map;
  Tree nt = Adapt(tree, old_flux, 0.05f, new_w, map);
  const int nn = nt.NumNodes();
  Stats ns(nn);
  std::vector<float> prior(nn);
  std::vector<double> rc(nn);
  for (int i = 0; i < nn; ++i) {
    const int o = map[i].idx;
    const float scale = map[i].fractional_size * map[i].fractional_size;
    ns.mean[i] = stats.mean[o];
    ns.m2[i] = stats.m2[o] * (double)scale;
    ns.n[i] = (int64_t)((float)stats.n[o] * scale);  // OnlineVariance::CountsMultiplied
    prior[i] = scale * prev.probs[o];
    rc[i] = (double)scale * old_rc[o];
  }
  Sampled dst;
  dst.tree = nt;
  dst.means.resize(nn);
  dst.stddev.resize(nn);
  dst.probs.resize(nn);
  for (int i = 0; i < nn; ++i) {
    dst.means[i] = rc[i] * ns.mean[i];
    dst.stddev[i] = rc[i] * ns.MeanErr(i, kSafeError, 2);
    const double mix = MixFactor(ns.n[i]);
    dst.probs[i] = (float)((double)prior[i] * mix + (dst.means[i] + dst.stddev[i]) * (1. - mix));
  }
  Propagate(nt, dst.probs);
  dst.prior = dst.probs;
  tree = std::move(nt);
  stats = Stats(nn);
  return dst;
}

Sampled Learned::Bake() const {
  const int nn = tree.NumNodes();
  const int64_t root_count = stats.n[tree.root];
  Sampled dst;
  dst.tree = tree;
  dst.means.resize(nn);
  dst.stddev.resize(nn);
  dst.probs.resize(nn);
  for (int i = 0; i < nn; ++i) {
    const double rc = root_count > 0 ? (double)stats.n[i] / (double)root_count : 0.;
    dst.means[i] = rc * stats.mean[i];
    dst.stddev[i] = rc * stats.MeanErr(i, kSafeError, 10);
    dst.probs[i] = (float)dst.means[i];
  }
  dst.prior = dst.probs;
  Propagate(tree, dst.probs);
  return dst;
}

}  // namespace quad
}  // namespace guide
}  // namespace ttgpu
