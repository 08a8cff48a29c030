// Binned-SAH BVH build on the host CPU (multi-threaded for the top levels).
#include "bvh_build.hpp"
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <thread>

namespace ttgpu {
namespace {

constexpr int kBins = 32;
constexpr float kInf = std::numeric_limits<float>::infinity();

struct Box {
  float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
  void grow(const float *l, const float *h) {
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], l[k]); hi[k] = std::max(hi[k], h[k]); }
  }
  void grow(const Box &b) { grow(b.lo, b.hi); }
  float area() const {
    float d[3];
    for (int k = 0; k < 3; ++k) d[k] = std::max(0.f, hi[k] - lo[k]);
    return 2.f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
  }
};

struct TmpNode {
  Box box;
  int left = -1, right = -1;  // tmp indices (inner)
  int start = 0, count = 0;   // leaf range
  int depth = 0;
};

struct Builder {
  const std::vector<BuildPrim> &prims;
  std::vector<float> cen;
  std::vector<uint32_t> idx;
  std::vector<TmpNode> tmp;
  std::atomic<int> next{0};
  int max_leaf;
  std::atomic<int> max_depth{0};

  Builder(const std::vector<BuildPrim> &p, int ml) : prims(p), max_leaf(ml) {
    size_t n = p.size();
    cen.resize(3 * n);
    idx.resize(n);
    for (size_t i = 0; i < n; ++i) {
      idx[i] = (uint32_t)i;
      for (int k = 0; k < 3; ++k) cen[3 * i + k] = 0.5f * (p[i].lo[k] + p[i].hi[k]);
    }
    tmp.resize(std::max<size_t>(1, 2 * n));
  }

  int Build(int begin, int end, int depth, int parallel_depth) {
    int ni = next.fetch_add(1);
    TmpNode &nd = tmp[ni];
    nd.depth = depth;
    int md = max_depth.load();
    while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {}
    Box cb;
    for (int i = begin; i < end; ++i) {
      const BuildPrim &p = prims[idx[i]];
      nd.box.grow(p.lo, p.hi);
      const float *c = &cen[3 * idx[i]];
      cb.grow(c, c);
    }
    const int n = end - begin;
    if (n <= 1) {
      nd.start = begin;
      nd.count = n;
      return ni;
    }
    // Binned SAH over all three axes.
    float best_cost = kInf;
    int best_axis = -1, best_split = -1;
    for (int axis = 0; axis < 3; ++axis) {
      float ext = cb.hi[axis] - cb.lo[axis];
      if (!(ext > 0.f)) continue;
      Box bins[kBins];
      int counts[kBins] = {0};
      float scale = kBins / ext;
      for (int i = begin; i < end; ++i) {
        int b = std::min(kBins - 1, (int)((cen[3 * idx[i] + axis] - cb.lo[axis]) * scale));
        counts[b]++;
        bins[b].grow(prims[idx[i]].lo, prims[idx[i]].hi);
      }
      float right_area[kBins];
      int right_count[kBins];
      Box acc;
      int cnt = 0;
      for (int b = kBins - 1; b > 0; --b) {
        acc.grow(bins[b]);
        cnt += counts[b];
        right_area[b] = acc.area();
        right_count[b] = cnt;
      }
      Box lacc;
      int lcnt = 0;
      for (int b = 0; b < kBins - 1; ++b) {
        lacc.grow(bins[b]);
        lcnt += counts[b];
        if (lcnt == 0 || right_count[b + 1] == 0) continue;
        float cost = lacc.area() * lcnt + right_area[b + 1] * right_count[b + 1];
        if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = b; }
      }
    }
    const float parent_area = std::max(nd.box.area(), 1e-30f);
    // Traversal : intersection cost ratio 1 : 1 (costs relative to one primitive test).
    float split_cost = 1.0f + best_cost / parent_area;
    float leaf_cost = (float)n;
    int mid;
    if (best_axis < 0) {
      if (n <= max_leaf) { nd.start = begin; nd.count = n; return ni; }
      mid = begin + n / 2;  // all centroids coincide: split the range
    } else {
      if (n <= max_leaf && leaf_cost <= split_cost) { nd.start = begin; nd.count = n; return ni; }
      float scale = kBins / (cb.hi[best_axis] - cb.lo[best_axis]);
      auto it = std::partition(idx.begin() + begin, idx.begin() + end, [&](uint32_t p) {
        int b = std::min(kBins - 1, (int)((cen[3 * p + best_axis] - cb.lo[best_axis]) * scale));
        return b <= best_split;
      });
      mid = (int)(it - idx.begin());
      if (mid == begin || mid == end) mid = begin + n / 2;
    }
    int l, r;
    if (parallel_depth > 0 && n > 65536) {
      std::thread t([&] { l = Build(begin, mid, depth + 1, parallel_depth - 1); });
      r = Build(mid, end, depth + 1, parallel_depth - 1);
      t.join();
    } else {
      l = Build(begin, mid, depth + 1, 0);
      r = Build(mid, end, depth + 1, 0);
    }
    tmp[ni].left = l;
    tmp[ni].right = r;
    return ni;
  }
};

void WriteChild(const TmpNode &c, int inner_index, float *xy, float *z, int *ch, int *cnt) {
  if (c.count > 0 || c.left < 0) {
    *ch = ~c.start;
    *cnt = c.count;
  } else {
    *ch = inner_index;
    *cnt = 0;
  }
  xy[0] = c.box.lo[0]; xy[1] = c.box.hi[0]; xy[2] = c.box.lo[1]; xy[3] = c.box.hi[1];
  z[0] = c.box.lo[2]; z[1] = c.box.hi[2];
}

}  // namespace

BuiltBvh BuildBvh(const std::vector<BuildPrim> &prims, int max_leaf, int num_threads) {
  BuiltBvh out;
  if (prims.empty()) return out;
  Builder b(prims, max_leaf);
  int pdepth = 0;
  unsigned hc = num_threads > 0 ? (unsigned)num_threads : std::max(1u, std::thread::hardware_concurrency());
  while ((1u << pdepth) < hc && pdepth < 6) ++pdepth;
  int root = b.Build(0, (int)prims.size(), 0, pdepth);
  int used = b.next.load();
  // Number the inner nodes in depth-first order (root first) for the device layout.
  std::vector<int> inner_id(used, -1);
  std::vector<int> order_nodes;
  std::vector<int> stack{root};
  while (!stack.empty()) {
    int t = stack.back();
    stack.pop_back();
    const TmpNode &nd = b.tmp[t];
    if (nd.left < 0) continue;
    inner_id[t] = (int)order_nodes.size();
    order_nodes.push_back(t);
    stack.push_back(nd.right);
    stack.push_back(nd.left);
  }
  const TmpNode &rn = b.tmp[root];
  for (int k = 0; k < 3; ++k) { out.root_lo[k] = rn.box.lo[k]; out.root_hi[k] = rn.box.hi[k]; }
  out.max_depth = b.max_depth.load();
  out.order = b.idx;
  if (order_nodes.empty()) {
    // Single leaf: a root whose right child is an empty leaf with an inverted box.
    out.nodes.assign(16, 0.f);
    float *n = out.nodes.data();
    int ch[4];
    WriteChild(rn, 0, n + 0, n + 8, &ch[0], &ch[2]);
    n[4] = kInf; n[5] = -kInf; n[6] = kInf; n[7] = -kInf; n[10] = kInf; n[11] = -kInf;
    ch[1] = ~0; ch[3] = 0;
    std::memcpy(n + 12, ch, sizeof(ch));
    out.num_nodes = 1;
    return out;
  }
  out.num_nodes = (int)order_nodes.size();
  out.nodes.assign(16 * (size_t)out.num_nodes, 0.f);
  for (int i = 0; i < out.num_nodes; ++i) {
    const TmpNode &nd = b.tmp[order_nodes[i]];
    const TmpNode &l = b.tmp[nd.left], &r = b.tmp[nd.right];
    float *n = &out.nodes[16 * (size_t)i];
    int ch[4];
    WriteChild(l, inner_id[nd.left], n + 0, n + 8, &ch[0], &ch[2]);
    WriteChild(r, inner_id[nd.right], n + 4, n + 10, &ch[1], &ch[3]);
    std::memcpy(n + 12, ch, sizeof(ch));
  }
  return out;
}

namespace {

// One child entry of a BVH2 node in the device layout (16 words: lxy, rxy, lrz, ch).
struct Child2 {
  float lo[3], hi[3];
  int ref;    // >= 0: BVH2 inner node, < 0: leaf ~first primitive
  int count;  // leaf primitives
  float Area() const {
    float d[3];
    for (int k = 0; k < 3; ++k) d[k] = std::max(0.f, hi[k] - lo[k]);
    return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
  }
};

Child2 ChildOf(const BuiltBvh &b, int node, int side) {
  const float *n = &b.nodes[16 * (size_t)node];
  const float *xy = n + 4 * side;
  const float *z = n + 8 + 2 * side;
  int ch[4];
  std::memcpy(ch, n + 12, sizeof(ch));
  Child2 c;
  c.lo[0] = xy[0]; c.hi[0] = xy[1]; c.lo[1] = xy[2]; c.hi[1] = xy[3]; c.lo[2] = z[0]; c.hi[2] = z[1];
  c.ref = ch[side];
  c.count = c.ref < 0 ? ch[2 + side] : 0;
  return c;
}

struct Collapser {
  const BuiltBvh &b2;
  BuiltBvh4 out;
  int Emit(int n2, int depth) {
    out.max_depth = std::max(out.max_depth, depth);
    std::vector<Child2> ch{ChildOf(b2, n2, 0), ChildOf(b2, n2, 1)};
    while (ch.size() < 4) {
      int best = -1;
      float best_area = -1.f;
      for (int i = 0; i < (int)ch.size(); ++i)
        if (ch[i].ref >= 0 && ch[i].Area() > best_area) { best = i; best_area = ch[i].Area(); }
      if (best < 0) break;
      const int r = ch[best].ref;
      ch.erase(ch.begin() + best);
      ch.push_back(ChildOf(b2, r, 0));
      ch.push_back(ChildOf(b2, r, 1));
    }
    const int id = out.num_nodes++;
    out.nodes.resize(32 * (size_t)out.num_nodes, 0.f);
    int child[4], count[4];
    float box[6][4];
    for (int s = 0; s < 4; ++s) {
      if (s < (int)ch.size()) {
        for (int k = 0; k < 3; ++k) { box[2 * k][s] = ch[s].lo[k]; box[2 * k + 1][s] = ch[s].hi[k]; }
        child[s] = ch[s].ref >= 0 ? Emit(ch[s].ref, depth + 1) : ch[s].ref;
        count[s] = ch[s].count;
      } else {
        // Empty slot: a box at +infinity misses every ray.
        for (int k = 0; k < 6; ++k) box[k][s] = kInf;
        child[s] = ~0;
        count[s] = 0;
      }
    }
    float *n = &out.nodes[32 * (size_t)id];
    // Rows: lo.x, hi.x, lo.y, hi.y, lo.z, hi.z (4 children each), child indices, leaf counts.
    for (int k = 0; k < 6; ++k) std::memcpy(n + 4 * k, box[k], sizeof(box[k]));
    std::memcpy(n + 24, child, sizeof(child));
    std::memcpy(n + 28, count, sizeof(count));
    return id;
  }
};

}  // namespace

BuiltBvh4 CollapseToBvh4(const BuiltBvh &b2) {
  Collapser c{b2, {}};
  if (b2.num_nodes > 0) c.Emit(0, 0);
  return c.out;
}

namespace {

// The device's decode of a quantized plane: q * 2^e is exact, so fma and multiply-add agree.
float DecodePlane(int q, float step, float origin) { return std::fma((float)q, step, origin); }

}  // namespace

std::vector<uint32_t> QuantizeBvh4(const BuiltBvh4 &b4) {
  std::vector<uint32_t> out(16 * (size_t)b4.num_nodes, 0u);
  for (int i = 0; i < b4.num_nodes; ++i) {
    const float *n = &b4.nodes[32 * (size_t)i];
    int child[4], count[4];
    std::memcpy(child, n + 24, sizeof(child));
    std::memcpy(count, n + 28, sizeof(count));
    bool valid[4];
    for (int c = 0; c < 4; ++c) valid[c] = !(child[c] < 0 && count[c] == 0);
    uint32_t *o = &out[16 * (size_t)i];
    uint32_t rows[6] = {0, 0, 0, 0, 0, 0};
    uint32_t exps = 0;
    for (int k = 0; k < 3; ++k) {
      const float *lo = n + 4 * (2 * k), *hi = n + 4 * (2 * k + 1);
      float org = kInf, top = -kInf;
      for (int c = 0; c < 4; ++c)
        if (valid[c]) { org = std::min(org, lo[c]); top = std::max(top, hi[c]); }
      if (!(org <= top)) org = top = 0.f;  // no child at all (empty tree)
      const double ext = (double)top - (double)org;
      int e = -100;
      if (ext > 0.) e = std::max(e, (int)std::ceil(std::log2(ext / 255.0)));
      for (;; ++e) {
        if (e > 127) throw std::runtime_error("BVH node extent too large for 8-bit planes");
        const float step = std::ldexp(1.f, e);
        bool ok = true;
        uint32_t rlo = 0, rhi = 0;
        for (int c = 0; c < 4 && ok; ++c) {
          if (!valid[c]) continue;
          int ql = (int)std::max(0.0, std::min(255.0, std::floor(((double)lo[c] - org) / step)));
          while (ql > 0 && DecodePlane(ql, step, org) > lo[c]) --ql;
          int qh = (int)std::max(0.0, std::min(255.0, std::ceil(((double)hi[c] - org) / step)));
          while (qh < 255 && DecodePlane(qh, step, org) < hi[c]) ++qh;
          ok = DecodePlane(ql, step, org) <= lo[c] && DecodePlane(qh, step, org) >= hi[c];
          rlo |= (uint32_t)ql << (8 * c);
          rhi |= (uint32_t)qh << (8 * c);
        }
        if (!ok) continue;
        rows[2 * k] = rlo;
        rows[2 * k + 1] = rhi;
        break;
      }
      std::memcpy(&o[k], &org, 4);
      exps |= (uint32_t)(e + 127) << (8 * k);
    }
    uint32_t counts = 0;
    for (int c = 0; c < 4; ++c) {
      if (count[c] < 0 || count[c] > 255) throw std::runtime_error("BVH leaf with more than 255 primitives");
      counts |= (uint32_t)count[c] << (8 * c);
    }
    o[3] = exps;
    for (int r = 0; r < 4; ++r) o[4 + r] = rows[r];
    o[8] = rows[4];
    o[9] = rows[5];
    o[10] = counts;
    o[11] = 0;
    std::memcpy(o + 12, child, sizeof(child));
  }
  return out;
}

}  // namespace ttgpu
