// Guided sampling from a cell's quadtree directional distribution on the GPU.
//
// The sphere maps to the unit square by v = (z + 1) / 2, u = (atan2(y, x) / pi + 1) / 2, which
// preserves area up to the factor 4 pi (src/path_guiding.cxx:112-129).  A point is drawn by
// hierarchical sample warping: per level the column is chosen from the marginal of the left /
// right halves and the row from the conditional in that column, and each random number is
// rescaled into the chosen interval (Clarberg et al. 2005, section 5.1, as in
// src/path_guiding_quadtree.cxx:168-244).  The density of a point is w_leaf / (w_root area) on
// the square (src/path_guiding_quadtree.cxx:247-259).  A descent reads one 32-byte node per level
// and stops after at most kQuadMaxDepth levels.
#pragma once
#include "device_math.hpp"
#include "device_scene.hpp"

namespace ttgpu {

struct QuadGuide {
  const QuadNode *nodes;
  int root;
  float w_root;
};

TT_HD QuadGuide quad_guide(const DeviceScene &sc, int cell) {
  const QuadCell c = sc.guide_qcells[cell];
  return QuadGuide{sc.guide_qnodes, c.root, c.w_root};
}

TT_HD float quad_pick(float4 v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w)); }
TT_HD int quad_pick(int4 v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w)); }

TT_HD float2 quad_sphere_to_square(float3 d) {
  return make_float2((atan2f(d.y, d.x) / kPi + 1.f) * 0.5f, (d.z + 1.f) * 0.5f);
}

// rho = sqrt(1 - z^2) = 2 sqrt(v (1 - v)) avoids the cancellation in 1 - z^2 near the poles, where
// z = 2v - 1 rounds to a coarse float; sin(pi x) / cos(pi x) avoid rounding pi x.
TT_HD float3 quad_square_to_sphere(float u, float v) {
  const float z = v * 2.f - 1.f;
  const float rho = 2.f * sqrtf(fmaxf(0.f, v * (1.f - v)));
  float sn, cs;
  sincospif(u * 2.f - 1.f, &sn, &cs);
  return normalize(mk(rho * cs, rho * sn, z));
}

// Density over the unit square.
TT_HD float quad_pdf_square(const QuadGuide &g, float2 p) {
  int node = g.root;
  float lo_u = 0.f, lo_v = 0.f, size = 1.f, w = g.w_root;
  for (int d = 0; d < kQuadMaxDepth && node >= 0; ++d) {
    const QuadNode n = g.nodes[node];
    size *= 0.5f;
    const int c = (p.x < lo_u + size ? 0 : 1) | (p.y < lo_v + size ? 0 : 2);
    if (c & 1) lo_u += size;
    if (c & 2) lo_v += size;
    w = quad_pick(n.w, c);
    node = quad_pick(n.child, c);
  }
  return w / (g.w_root * (size * size));
}

TT_HD float guide_pdf(const QuadGuide &g, float3 x) {
  return quad_pdf_square(g, quad_sphere_to_square(x)) * (1.f / (4.f * kPi));
}

// Warps (r0, r1) into the square; returns the point (density through quad_pdf_square).
TT_HD float2 quad_warp(const QuadGuide &g, float r0, float r1) {
  constexpr float kLessThan1 = 0x1.fffffep-1f;
  constexpr float kEpsF = 0x1.0p-23f;  // FLT_EPSILON
  r0 = fminf(r0, kLessThan1);
  r1 = fminf(r1, kLessThan1);
  int node = g.root;
  float lo_u = 0.f, lo_v = 0.f, size = 1.f;
  for (int d = 0; d < kQuadMaxDepth && node >= 0; ++d) {
    const QuadNode n = g.nodes[node];
    const float wl = n.w.x + n.w.z, wr = n.w.y + n.w.w;
    const float s0 = r0 * (wl + wr);
    int c = s0 < wl ? 0 : 1;
    r0 = s0 < wl ? s0 / (wl + kEpsF) : (s0 - wl) / (wr + kEpsF);
    const float wb = c ? n.w.y : n.w.x, wt = c ? n.w.w : n.w.z;
    const float s1 = r1 * (wb + wt);
    c |= s1 < wb ? 0 : 2;
    r1 = s1 < wb ? s1 / (wb + kEpsF) : (s1 - wb) / (wt + kEpsF);
    r0 = fminf(r0, kLessThan1);
    r1 = fminf(r1, kLessThan1);
    size *= 0.5f;
    if (c & 1) lo_u += size;
    if (c & 2) lo_v += size;
    node = quad_pick(n.child, c);
  }
  const float u = fmaxf(fminf(size * r0 + lo_u, (lo_u + size) * kLessThan1), lo_u);
  const float v = fmaxf(fminf(size * r1 + lo_v, (lo_v + size) * kLessThan1), lo_v);
  return make_float2(u, v);
}

TT_HD float3 guide_sample(const QuadGuide &g, float u0, float u1, float /*u2*/) {
  const float2 p = quad_warp(g, u0, u1);
  return quad_square_to_sphere(p.x, p.y);
}

}  // namespace ttgpu
