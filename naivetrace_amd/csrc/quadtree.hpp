// Quadtree directional distribution of the guided integrator (host side): the
// tree over the equal-area square [0,1]^2 of the sphere, its builder and
// adaptor, the per-node online statistics, and the learned / sampled radiance
// distributions with the UCB-style node sampling weights.
//
// Follows ToyTrace:
//   src/path_guiding_quadtree.hxx:20-352, .cxx:32-272   tree, region, builder,
//       adaptor, PushWeight, PropagateLeafWeightsToParents, Sample (Clarberg et
//       al. 2005, section 5.1), Pdf
//   src/path_guiding.hxx:190-340, src/path_guiding.cxx:108-294   sphere <->
//       square map, stochastic direction filter, IncrementalFit, InitialFit,
//       IterationUpdateAndBake, Bake
//   src/sampler.hxx:248-270, 307-520   Welford statistics (SoaOnlineVariance)
// Trees are numbered in depth-first pre-order (parent, then children 0..3),
// the order in which the reference allocates them; child 0 is the bottom-left
// quadrant, 1 bottom-right, 2 top-left, 3 top-right.
#pragma once
#include <array>
#include <cstdint>
#include <utility>
#include <vector>

namespace ttgpu {
namespace guide {
namespace quad {

constexpr int kMaxDepth = 8;                  // path_guiding_quadtree.hxx:22
constexpr int kMaxNodes = (1 << 16) - 2;      // path_guiding_quadtree.hxx:21
constexpr double kJacobianInv = 1. / (4. * 3.14159265358979323846);
// LargeFloat = FLT_MAX / 16 (vec3f.hxx:23) and safe_error_number (path_guiding.hxx:199).
constexpr double kLargeFloat = 3.4028234663852886e38 / 16.;
constexpr double kSafeError = 0.01 * kLargeFloat * (1. / 65536.);  // 0.25^8

struct Point {
  float u, v;
};

struct Tree {
  std::vector<std::array<int32_t, 4>> child;  // child[n][0] < 0: n is a leaf
  int root = 0;
  Tree() : child{{-1, -1, -1, -1}} {}
  int NumNodes() const { return (int)child.size(); }
  bool IsLeaf(int n) const { return child[n][0] < 0; }
};

// Descent through the dyadic regions: child number (x >= centre) | 2 (y >= centre).
struct Descent {
  Point p;
  float lo[2] = {0.f, 0.f}, size = 1.f;
  int node;
  int depth = 0;  // number of steps taken
  Descent(Point q, const Tree &t) : p(q), node(t.root) {}
  bool Leaf(const Tree &t) const { return t.IsLeaf(node); }
  void Step(const Tree &t);
};

// MapSphereToTree / MapTreeToSphere (path_guiding.cxx:112-129): v = (z + 1) / 2,
// u = (atan2(y, x) / pi + 1) / 2; equal-area, so densities differ by 1 / (4 pi).
Point SphereToTree(const float d[3]);
void TreeToSphere(Point p, float out[3]);

// Builder (path_guiding_quadtree.hxx:317-352, .cxx:32-88): a node with weight sum above
// fraction x total and depth < kMaxDepth (the root has depth 1) is split at its centre.
Tree Build(const std::vector<Point> &pts, const std::vector<float> &w, float fraction, std::vector<float> &node_w);

// TreeAdaptor (path_guiding_quadtree.hxx:209-315): threshold = fraction x weight of the root.
struct NewToOld {
  int idx;
  float fractional_size;  // edge length relative to the old node (< 1 where a leaf was split)
};
Tree Adapt(const Tree &t, const std::vector<float> &w, float fraction, std::vector<float> &new_w,
           std::vector<NewToOld> &mapping);

// Sums leaf weights into every branch (PropagateLeafWeightsToParents).
void Propagate(const Tree &t, std::vector<float> &w);
// Hierarchical sample warping with rescaled random numbers; returns the point and its density
// on the square (w_leaf / (w_root area)).
std::pair<Point, float> Sample(const Tree &t, const std::vector<float> &w, float r0, float r1);
float Pdf(const Tree &t, const std::vector<float> &w, Point p);

// Per-node Welford statistics (SoaOnlineVariance<double, long>).
struct Stats {
  std::vector<double> mean, m2;
  std::vector<int64_t> n;
  explicit Stats(int size = 1) : mean(size, 0.), m2(size, 0.), n(size, 0) {}
  void Add(int i, double x);
  // sqrt(M2) / (n - 1) where n >= min_count, else fill (SoaOnlineVariance::MeanErr).
  double MeanErr(int i, double fill, int64_t min_count) const;
};

// RadianceDistributionSampled: what the renderer samples.
struct Sampled {
  Tree tree;
  std::vector<double> means{1.}, stddev{0.};
  std::vector<float> probs{1.f}, prior{1.f};
  double Pdf(const float dir[3]) const;
  // Leaf depth under the direction (number of descents), for the direction filter.
  int LeafDepth(const float dir[3]) const;
  // ComputeStochasticFilteredDirection (path_guiding.cxx:132-153): a uniform direction in the
  // cone of half-angle 0.5^(depth + 1) sqrt(4 pi) around dir.
  void FilterDirection(const float dir[3], double u0, double u1, float out[3]) const;
};

struct Sample1 {
  const float *dir;
  float weight;
  bool original;
};

// RadianceDistributionLearned.
struct Learned {
  Tree tree;
  Stats stats;
  // Builds the tree from every sample (fraction 0.1) and pushes them all.
  void InitialFit(const std::vector<Sample1> &s);
  // Pushes the filtered (non-original) samples and sets dst's node probabilities: prior x mix +
  // (means + stddev) (1 - mix) + root flux sqrt(log(round) / (count + 1)), mix = 100 / (100 +
  // max(count - 2, 0)), then propagated.  Without any pushed sample in the tree, dst is kept.
  void IncrementalFit(const std::vector<Sample1> &s, int round, Sampled &dst);
  // Adapts the tree (fraction 0.05 of the root flux), carries the statistics over with counts
  // scaled by the squared fractional size, bakes and resets the statistics; returns `prev`
  // unchanged when the root has 10 samples or fewer.
  Sampled IterationUpdateAndBake(const Sampled &prev);
  Sampled Bake() const;
};

}  // namespace quad
}  // namespace guide
}  // namespace ttgpu
