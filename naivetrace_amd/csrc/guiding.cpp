// Host recorder of the guided integrator: kd-tree, vMF-mixture EM, rounds.
// See guiding.hpp for the reference files followed.
#include "guiding.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>
#include <thread>

namespace ttgpu {
namespace guide {

namespace {
constexpr double kPi = 3.14159265358979323846;

void FrameZ(const float z[3], float x[3], float y[3]) {  // Duff et al. 2017
  const float sign = std::copysign(1.f, z[2]);
  const float a = -1.f / (sign + z[2]);
  const float b = z[0] * z[1] * a;
  x[0] = 1.f + sign * z[0] * z[0] * a; x[1] = sign * b; x[2] = -sign * z[0];
  y[0] = b; y[1] = sign + z[1] * z[1] * a; y[2] = -z[1];
}
}  // namespace

void InitUnitSphere(Vmf &m) {
  // Lobe directions of the fixed 8-lobe prior (distribution_mixture_models.cxx:171-179).
  static const float kMeans[kLobes][3] = {
      {0.66778004f, 0.73539025f, 0.11519922f},  {-0.9836298f, -0.008926501f, 0.1799797f},
      {-0.14629413f, -0.7809748f, 0.60718733f}, {0.49043104f, 0.04368747f, -0.87038434f},
      {-0.10088497f, 0.42765132f, 0.8982965f},  {-0.35999632f, -0.6945797f, -0.62286574f},
      {-0.43178025f, 0.7588791f, -0.48751223f}, {0.860208f, -0.47938251f, 0.17388113f}};
  for (int k = 0; k < kLobes; ++k) {
    for (int c = 0; c < 3; ++c) m.mu[k][c] = kMeans[k][c];
    m.kappa[k] = 2.f;
    m.w[k] = 1.f / kLobes;
  }
}

float ComponentPdf(const float mu[3], float k, const float x[3]) {
  const float d = mu[0] * x[0] + mu[1] * x[1] + mu[2] * x[2];
  return k / (float(2. * kPi) * (1.f - std::exp(-2.f * k))) * std::exp(k * (d - 1.f));
}

float Pdf(const Vmf &m, const float x[3]) {
  float p = 0.f;
  for (int k = 0; k < kLobes; ++k) p += m.w[k] * ComponentPdf(m.mu[k], m.kappa[k], x);
  return p;
}

void Sample(const Vmf &m, double r0, double r1, double r2, float out[3]) {
  int idx = kLobes - 1;
  double acc = 0.;
  for (int k = 0; k < kLobes; ++k) {
    acc += m.w[k];
    if (r0 < acc) { idx = k; break; }
  }
  const float k = m.kappa[idx];
  const float vx = std::cos(float(2. * kPi) * (float)r1), vy = std::sin(float(2. * kPi) * (float)r1);
  const float w = 1.f + std::log((float)r2 + (1.f - (float)r2) * std::exp(-2.f * k)) / k;
  const float rho = std::sqrt(std::max(0.f, 1.f - w * w));
  float x[3], y[3];
  FrameZ(m.mu[idx], x, y);
  for (int c = 0; c < 3; ++c) out[c] = x[c] * rho * vx + y[c] * rho * vy + m.mu[idx][c] * w;
}

float MeanCosineToConc(float r) {  // distribution_mixture_models.hxx:46-54
  const float t = 1.f - 1.e-6f;
  if (r >= t) r = t;
  return r * (3.f - r * r) / (1.f - r * r);
}

void FitSample(Vmf &m, FitData &d, const FitParams &p, const float x[3], float weight) {
  // UpdateStatistics (distribution_mixture_models.cxx:270-300, the running-mean branch).
  double r[kLobes], rs = 0.;
  for (int k = 0; k < kLobes; ++k) {
    r[k] = (double)m.w[k] * ComponentPdf(m.mu[k], m.kappa[k], x) + kEps;
    rs += r[k];
  }
  ++d.weight_count;
  d.weight_mean += (weight - d.weight_mean) / (double)d.weight_count;
  ++d.count;
  const double inv_n = 1. / (double)d.count;
  for (int k = 0; k < kLobes; ++k) {
    const double rk = r[k] / rs;
    d.resp[k] += (weight * rk - d.resp[k]) * inv_n;
    for (int c = 0; c < 3; ++c) d.pos[k][c] += (weight * rk * x[c] - d.pos[k][c]) * inv_n;
  }
  if (d.count % p.m_step_every == 0) {
    MaximizationStep(m, d, p);
    const double wm = d.weight_mean;
    const int64_t wc = d.weight_count;
    d = FitData{};
    d.weight_mean = wm;
    d.weight_count = wc;
  }
}

void MaximizationStep(Vmf &m, const FitData &d, const FitParams &p) {
  // distribution_mixture_models.cxx:303-343
  const double n = (double)d.count;
  const Vmf &prior = *p.prior;
  double wsum = 0.;
  for (int k = 0; k < kLobes; ++k) {
    const double scaled = d.resp[k] / (d.weight_mean + kEps);
    m.w[k] = (float)((p.prior_nu - 1.) / n * prior.w[k] + scaled + kEps);
    const double dim = d.weight_mean * p.prior_tau / n;
    double mu[3], norm2 = 0.;
    for (int c = 0; c < 3; ++c) {
      mu[c] = (dim * prior.mu[k][c] + d.pos[k][c]) / (dim + d.resp[k] + kEps);
      norm2 += mu[c] * mu[c];
    }
    const double norm = std::sqrt(norm2);
    for (int c = 0; c < 3; ++c) m.mu[k][c] = norm > kEps ? (float)(mu[c] / norm) : (c == 0 ? 1.f : 0.f);
    const double pos_norm = std::sqrt(d.pos[k][0] * d.pos[k][0] + d.pos[k][1] * d.pos[k][1] + d.pos[k][2] * d.pos[k][2]);
    const float mean_cosine = (float)(pos_norm / (d.resp[k] + kEps));
    const float conc = MeanCosineToConc(mean_cosine);
    const float dim_alpha = (float)(p.prior_alpha / n);
    m.kappa[k] = std::min(kKappaMax, std::max(kKappaMin, (prior.kappa[k] * dim_alpha + conc) / (dim_alpha + 1.f)));
    wsum += m.w[k];
  }
  for (int k = 0; k < kLobes; ++k) m.w[k] = (float)(m.w[k] / wsum);
}

void OnlineCov3::Add(const double p[3]) {
  ++n;
  double dlt[3];
  for (int a = 0; a < 3; ++a) {
    dlt[a] = p[a] - mean[a];
    mean[a] += dlt[a] / (double)n;
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) c[a][b] += dlt[a] * (p[b] - mean[b]);
}
void OnlineCov3::Cov(double out[3][3]) const {
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) out[a][b] = n > 1 ? c[a][b] / (double)(n - 1) : 0.;
}
void OnlineCov3::Var(double out[3]) const {
  for (int a = 0; a < 3; ++a) out[a] = n > 1 ? c[a][a] / (double)(n - 1) : 0.;
}

int KdTree::Lookup(const double p[3]) const {
  int h = root;
  while (h >= 0) {
    const Node &nd = nodes[h];
    h = p[nd.axis] < nd.pos ? nd.left : nd.right;
  }
  return ~h;
}

namespace {
struct Builder {
  const std::vector<std::array<double, 3>> &pts;
  std::vector<int> &order;
  std::vector<std::pair<int, int>> &ranges;
  KdTree &t;
  int min_points, max_depth;
  int Leaf(int b, int e) {
    ranges.push_back({b, e});
    return ~(t.num_leafs++);
  }
  int Build(int b, int e, int depth) {
    const int n = e - b;
    if (n < 2 * min_points || depth >= max_depth) return Leaf(b, e);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = b; i < e; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], pts[order[i]][a]);
        hi[a] = std::max(hi[a], pts[order[i]][a]);
      }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
      if (hi[a] - lo[a] > hi[axis] - lo[axis]) axis = a;
    const int half = n / 2;
    std::nth_element(order.begin() + b, order.begin() + b + half, order.begin() + e,
                     [&](int x, int y) { return pts[x][axis] < pts[y][axis]; });
    const double pos = pts[order[b + half]][axis];
    if (half < min_points || n - half < min_points) return Leaf(b, e);
    const int left = Build(b, b + half, depth + 1);
    const int right = Build(b + half, e, depth + 1);
    t.nodes.push_back({axis, pos, left, right});
    return (int)t.nodes.size() - 1;
  }
};

struct Adaptor {
  const KdTree &old;
  const std::function<std::pair<int, double>(int)> &decide;
  std::vector<std::pair<int, int>> &mapping;
  KdTree t;
  int Adapt(int h, int depth) {
    if (h < 0) {
      const int leaf = ~h;
      auto [axis, pos] = decide(leaf);
      if (axis >= 0 && depth < kMaxTreeDepth) {
        const int l = t.num_leafs++, r = t.num_leafs++;
        mapping[leaf] = {l, r};
        t.nodes.push_back({axis, pos, ~l, ~r});
        return (int)t.nodes.size() - 1;
      }
      const int c = t.num_leafs++;
      mapping[leaf] = {c, -1};
      return ~c;
    }
    const KdTree::Node nd = old.nodes[h];
    const int l = Adapt(nd.left, depth + 1);
    const int r = Adapt(nd.right, depth + 1);
    t.nodes.push_back({nd.axis, nd.pos, l, r});
    return (int)t.nodes.size() - 1;
  }
};

void LeafBoxesRec(const KdTree &t, int h, const Box3 &b, std::vector<Box3> &out) {
  if (h < 0) {
    out[~h] = b;
    return;
  }
  const auto &nd = t.nodes[h];
  Box3 l = b, r = b;
  l.hi[nd.axis] = nd.pos;
  r.lo[nd.axis] = nd.pos;
  LeafBoxesRec(t, nd.left, l, out);
  LeafBoxesRec(t, nd.right, r, out);
}
}  // namespace

KdTree BuildTree(const std::vector<std::array<double, 3>> &pts, int min_points, int max_depth, std::vector<int> &order,
                 std::vector<std::pair<int, int>> &leaf_ranges) {
  KdTree t;
  t.num_leafs = 0;
  order.resize(pts.size());
  std::iota(order.begin(), order.end(), 0);
  leaf_ranges.clear();
  Builder b{pts, order, leaf_ranges, t, min_points, max_depth};
  t.root = b.Build(0, (int)pts.size(), 1);
  return t;
}

KdTree AdaptTree(const KdTree &t, const std::function<std::pair<int, double>(int)> &decide,
                 std::vector<std::pair<int, int>> &mapping) {
  mapping.assign(t.num_leafs, {-1, -1});
  Adaptor a{t, decide, mapping, KdTree{}};
  a.t.num_leafs = 0;
  a.t.root = a.Adapt(t.root, 1);
  return a.t;
}

void LeafBoxes(const KdTree &t, const Box3 &root_box, std::vector<Box3> &out) {
  out.resize(t.num_leafs);
  LeafBoxesRec(t, t.root, root_box, out);
}

void SymEig3(const double a_in[3][3], double val[3], double vec[3][3]) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      a[i][j] = a_in[i][j];
      vec[i][j] = i == j ? 1. : 0.;
    }
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (off < 1e-30) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (std::abs(a[p][q]) < 1e-300) continue;
        const double theta = (a[q][q] - a[p][p]) / (2. * a[p][q]);
        const double t = (theta >= 0 ? 1. : -1.) / (std::abs(theta) + std::sqrt(theta * theta + 1.));
        const double c = 1. / std::sqrt(t * t + 1.), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A = J^T A J
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = vec[k][p], vkq = vec[k][q];
          vec[k][p] = c * vkp - s * vkq;
          vec[k][q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 3; ++i) val[i] = a[i][i];
}

namespace {
// InitializePcaFrame (path_guiding.cxx:592-610): eigenvectors scaled by the standard deviations.
void PcaFrame(const OnlineCov3 &s, double frame[3][3]) {
  double cov[3][3], val[3], vec[3][3];
  s.Cov(cov);
  SymEig3(cov, val, vec);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) frame[i][j] = vec[i][j] * std::sqrt(std::max(val[j], 2.220446049250313e-16));
}
}  // namespace

Recorder::Recorder(int subdiv_factor, uint64_t seed, int threads, bool quadtree)
    : subdiv_factor_(std::max(1, subdiv_factor)), quadtree_(quadtree), seed_(seed),
      threads_(threads > 0 ? threads : (int)std::thread::hardware_concurrency()) {
  InitUnitSphere(prior_);
  tree_.root = ~0;
  tree_.num_leafs = 1;
  cells_.resize(1);
}

void Recorder::ParallelFor(int n, const std::function<void(int, int)> &fn) {
  const int nt = std::max(1, std::min(threads_, n));
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (int i = t; i < n; i += nt) fn(i, t);
    });
  for (auto &x : th) x.join();
}

FitParams Recorder::ParamsOf(const CellData &c) const {
  // RadianceDistributionLearned::Parameters (path_guiding.cxx:62-77)
  FitParams p;
  const double strength = std::max(1., 0.01 * (double)c.max_num_samples);
  p.prior_alpha = p.prior_nu = p.prior_tau = (float)strength;
  p.m_step_every = std::max(1, (int)strength * 10);
  p.prior = &prior_;
  return p;
}

void Recorder::FinalizeRound(std::vector<Record> &records) {
  ++round_;
  sub_round_ = 1;
  if (round_ <= 1) AdaptInitial(records);
  else FitTheSamples(records);
}

void Recorder::AdaptInitial(std::vector<Record> &records) {
  if (records.empty()) return;
  std::vector<std::array<double, 3>> pts(records.size());
  for (size_t i = 0; i < records.size(); ++i) pts[i] = {records[i].pos[0], records[i].pos[1], records[i].pos[2]};
  std::vector<int> order;
  std::vector<std::pair<int, int>> ranges;
  tree_ = BuildTree(pts, subdiv_factor_, kMaxTreeDepth, order, ranges);
  cells_.assign(tree_.num_leafs, CellData{});
  ParallelFor(tree_.num_leafs, [&](int i, int) {
    CellData &cd = cells_[i];
    const auto [b, e] = ranges[i];
    cd.last_num_samples = cd.max_num_samples = e - b;
    for (int k = b; k < e; ++k) cd.leaf_stats.Add(records[order[k]].pos);
    PcaFrame(cd.leaf_stats, cd.cov_frame);
    if (quadtree_) {
      // InitialFit on every record of the cell, then Bake (path_guiding.cxx:660-662).
      std::vector<quad::Sample1> s;
      s.reserve(e - b);
      for (int k = b; k < e; ++k) s.push_back({records[order[k]].dir, records[order[k]].weight, true});
      cd.qlearned.InitialFit(s);
      cd.qcurrent = cd.qlearned.Bake();
      return;
    }
    const FitParams p = ParamsOf(cd);
    for (int k = b; k < e; ++k) {
      const Record &r = records[order[k]];
      if (r.weight > 0.f) FitSample(cd.learned, cd.fit, p, r.dir, r.weight);
    }
    cd.current = cd.learned;
  });
  previous_max_samples_per_cell_ = subdiv_factor_;
  previous_total_samples_ = (int64_t)records.size();
}

void Recorder::FitTheSamples(std::vector<Record> &records) {
  const size_t n0 = records.size();
  if (n0 == 0) return;
  std::vector<int> cell(n0 * 2);
  for (size_t i = 0; i < n0; ++i) cell[i] = tree_.Lookup(records[i].pos);
  // Stochastic smoothing filter: a jittered copy of every record (path_guiding.cxx:353-428).
  records.resize(n0 * 2);
  const int nt = std::max(1, threads_);
  ParallelFor(nt, [&](int t, int) {
    std::mt19937_64 rng(seed_ * 0x9E3779B97F4A7C15ull + (uint64_t)round_ * 1315423911ull + (uint64_t)t);
    std::uniform_real_distribution<double> U(0., 1.);
    for (size_t i = t; i < n0; i += nt) {
      const Record &r = records[i];
      const CellData &cd = cells_[cell[i]];
      Record f = r;
      double rv[3] = {U(rng) * 2. - 1., U(rng) * 2. - 1., U(rng) * 2. - 1.};
      for (int a = 0; a < 3; ++a)
        f.pos[a] = r.pos[a] + 3. * (cd.cov_frame[a][0] * rv[0] + cd.cov_frame[a][1] * rv[1] + cd.cov_frame[a][2] * rv[2]);
      f.original = false;
      if (quadtree_) {
        // The quadtree's filter spreads over the leaf's angular footprint (path_guiding.cxx:132-153).
        const double u0 = U(rng), u1 = U(rng);
        cd.qcurrent.FilterDirection(r.dir, u0, u1, f.dir);
        records[n0 + i] = f;
        cell[n0 + i] = tree_.Lookup(f.pos);
        continue;
      }
      // ComputeStochasticFilteredDirection (path_guiding.cxx:80-88)
      const float pdf = Pdf(cd.current, r.dir);
      const float a = -0.1f / (pdf * float(2. * kPi) + 1.e-9f) + 1.f;
      const float cos_open = std::min(std::max(a, -1.f), 1.f);
      const double u0 = U(rng), u1 = U(rng);
      const double z = 1. - u1 * (1. - cos_open);
      const double s = std::sqrt(std::max(0., 1. - z * z));
      const double om = 2. * kPi * u0;
      float fx[3], fy[3];
      FrameZ(r.dir, fx, fy);
      for (int c = 0; c < 3; ++c) f.dir[c] = (float)(fx[c] * std::cos(om) * s + fy[c] * std::sin(om) * s + r.dir[c] * z);
      records[n0 + i] = f;
      cell[n0 + i] = tree_.Lookup(f.pos);
    }
  });
  // Sort into cells, shuffle, fit.
  std::vector<std::vector<int>> per_cell(cells_.size());
  for (size_t i = 0; i < records.size(); ++i) per_cell[cell[i]].push_back((int)i);
  ParallelFor((int)cells_.size(), [&](int c, int) {
    auto &idx = per_cell[c];
    if (idx.empty()) return;
    std::mt19937_64 rng(seed_ * 0x2545F4914F6CDD1Dull + (uint64_t)round_ * 7919ull + (uint64_t)c);
    std::shuffle(idx.begin(), idx.end(), rng);
    CellData &cd = cells_[c];
    if (quadtree_) {
      // IncrementalFit updates the sampled distribution of the cell in place (path_guiding.cxx:476-481).
      std::vector<quad::Sample1> s;
      s.reserve(idx.size());
      for (int i : idx) s.push_back({records[i].dir, records[i].weight, records[i].original});
      cd.qlearned.IncrementalFit(s, sub_round_, cd.qcurrent);
      for (int i : idx)
        if (records[i].original) cd.leaf_stats.Add(records[i].pos);
      return;
    }
    const FitParams p = ParamsOf(cd);
    for (int i : idx) {
      const Record &r = records[i];
      if (r.weight > 0.f) FitSample(cd.learned, cd.fit, p, r.dir, r.weight);
      if (r.original) cd.leaf_stats.Add(r.pos);
    }
  });
  ++sub_round_;
}

void Recorder::PrepareAdaptedStructures() {
  if (round_ > 1) AdaptIncremental();
}

void Recorder::AdaptIncremental() {
  // path_guiding.cxx:650-812
  int64_t num_fit = 0;
  for (const auto &c : cells_) num_fit += c.leaf_stats.n;
  const int64_t max_per_cell = num_fit
      ? (int64_t)(std::sqrt((double)num_fit) * (double)previous_max_samples_per_cell_ / std::sqrt((double)previous_total_samples_))
      : previous_max_samples_per_cell_;
  auto decide = [&](int c) -> std::pair<int, double> {
    const auto &s = cells_[c].leaf_stats;
    if (s.n > max_per_cell) {
      double var[3];
      s.Var(var);
      int axis = 0;
      for (int a = 1; a < 3; ++a)
        if (var[a] > var[axis]) axis = a;
      return {axis, s.mean[axis]};
    }
    return {-1, 0.};
  };
  std::vector<std::pair<int, int>> mapping;
  KdTree t = AdaptTree(tree_, decide, mapping);
  std::vector<CellData> nd(t.num_leafs);
  for (int i = 0; i < (int)cells_.size(); ++i) {
    CellData &src = cells_[i];
    quad::Sampled baked;
    if (quadtree_) baked = src.qlearned.IterationUpdateAndBake(src.qcurrent);
    for (int dst : {mapping[i].first, mapping[i].second}) {
      if (dst < 0) continue;
      CellData &d = nd[dst];
      if (quadtree_) {
        d.qcurrent = baked;
        d.qlearned = src.qlearned;
      }
      d.current = src.learned;  // IterationUpdateAndBake
      d.learned = src.learned;
      d.fit = src.fit;
      if (src.leaf_stats.n >= 10) PcaFrame(src.leaf_stats, d.cov_frame);
      else std::copy(&src.cov_frame[0][0], &src.cov_frame[0][0] + 9, &d.cov_frame[0][0]);
      d.last_num_samples = src.leaf_stats.n;
      d.max_num_samples = std::max(d.last_num_samples, src.max_num_samples);
    }
  }
  tree_ = std::move(t);
  cells_ = std::move(nd);
  previous_max_samples_per_cell_ = max_per_cell;
  previous_total_samples_ = num_fit;
}

}  // namespace guide
}  // namespace ttgpu
