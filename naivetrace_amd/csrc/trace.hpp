// BVH traversal on the GPU: closest hit and any hit, FP32, watertight
// ray/triangle test (Woop, Benthin & Wald 2013) and spheres.  Replaces
// rtcIntersect1 / rtcOccluded1 (src/embreeaccelerator.cxx:111,302).
//
// The traversal stack lives in LDS, one column per lane (bank-conflict free),
// with a small private overflow area for unusually deep trees.
#pragma once
#include "device_math.hpp"

namespace ttgpu {

constexpr int kTraceBlock = 128;    // threads per block of the trace kernels
constexpr int kLdsStack = 24;       // LDS stack entries per lane
constexpr int kSpillStack = 40;     // private overflow entries per lane

struct Ray {
  float3 org, dir;
  float tmax;
};

struct HitRec {
  float t, u, v;
  int prim;  // leaf-order primitive, -1 = none
};

struct RayPre {
  float3 inv, oinv;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

TT_HD RayPre precompute(const Ray &r) {
  RayPre p;
  const float tiny = 1e-30f;
  float dx = fabsf(r.dir.x) > tiny ? r.dir.x : copysignf(tiny, r.dir.x);
  float dy = fabsf(r.dir.y) > tiny ? r.dir.y : copysignf(tiny, r.dir.y);
  float dz = fabsf(r.dir.z) > tiny ? r.dir.z : copysignf(tiny, r.dir.z);
  p.inv = mk(1.f / dx, 1.f / dy, 1.f / dz);
  p.oinv = mk(r.org.x * p.inv.x, r.org.y * p.inv.y, r.org.z * p.inv.z);
  float ax = fabsf(r.dir.x), ay = fabsf(r.dir.y), az = fabsf(r.dir.z);
  int kz = 0;
  if (ay > ax) kz = 1;
  if (az > fmaxf(ax, ay)) kz = 2;
  int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
  float dkz = kz == 0 ? r.dir.x : (kz == 1 ? r.dir.y : r.dir.z);
  if (dkz < 0.f) { int t = kx; kx = ky; ky = t; }
  float dkx = kx == 0 ? r.dir.x : (kx == 1 ? r.dir.y : r.dir.z);
  float dky = ky == 0 ? r.dir.x : (ky == 1 ? r.dir.y : r.dir.z);
  p.kx = kx; p.ky = ky; p.kz = kz;
  p.Sx = dkx / dkz; p.Sy = dky / dkz; p.Sz = 1.f / dkz;
  return p;
}

TT_HD float comp(float3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Watertight test; on a hit within (0, tmax) writes t and barycentrics (b1, b2) of v1, v2.
TT_HD bool intersect_triangle(const RayPre &p, float3 org, float3 v0, float3 v1, float3 v2, float tmax, float &t,
                                                   float &b1, float &b2) {
  float3 A = v0 - org, B = v1 - org, C = v2 - org;
  const float Akz = comp(A, p.kz), Bkz = comp(B, p.kz), Ckz = comp(C, p.kz);
  const float Ax = comp(A, p.kx) - p.Sx * Akz, Ay = comp(A, p.ky) - p.Sy * Akz;
  const float Bx = comp(B, p.kx) - p.Sx * Bkz, By = comp(B, p.ky) - p.Sy * Bkz;
  const float Cx = comp(C, p.kx) - p.Sx * Ckz, Cy = comp(C, p.ky) - p.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
  const float det = U + V + W;
  if (det == 0.f) return false;
  const float Az = p.Sz * Akz, Bz = p.Sz * Bkz, Cz = p.Sz * Ckz;
  const float T = U * Az + V * Bz + W * Cz;
  const float inv = 1.f / det;
  const float tt = T * inv;
  if (!(tt > 0.f && tt < tmax)) return false;
  // Conservative lower bound on t (PBRT-v3 Triangle::Intersect): rays leaving a surface do not
  // re-hit it at a distance below the rounding error of t.
  {
    constexpr float e = kFltEps;
    const float g2 = 2.f * e / (1.f - 2.f * e), g3 = 3.f * e / (1.f - 3.f * e), g5 = 5.f * e / (1.f - 5.f * e);
    const float maxZt = fmaxf(fabsf(Az), fmaxf(fabsf(Bz), fabsf(Cz)));
    const float maxXt = fmaxf(fabsf(Ax), fmaxf(fabsf(Bx), fabsf(Cx)));
    const float maxYt = fmaxf(fabsf(Ay), fmaxf(fabsf(By), fabsf(Cy)));
    const float deltaZ = g3 * maxZt, deltaX = g5 * (maxXt + maxZt), deltaY = g5 * (maxYt + maxZt);
    const float deltaE = 2.f * (g2 * maxXt * maxYt + deltaY * maxXt + deltaX * maxYt);
    const float maxE = fmaxf(fabsf(U), fmaxf(fabsf(V), fabsf(W)));
    const float deltaT = 3.f * (g3 * maxE * maxZt + deltaE * maxZt + deltaZ * maxE) * fabsf(inv);
    if (tt <= deltaT) return false;
  }
  t = tt;
  b1 = V * inv;
  b2 = W * inv;
  return true;
}

// Sphere: nearest root in (eps, tmax), numerically stable form, in FP64 so that spheres of
// planetary size (thousands of scene units) keep their hit distances and inside/outside
// decisions exact to well below the self-intersection offset.  The lower bound eps is
// 1e-9 of the sphere's extent and absorbs the residual root of a ray leaving the surface.
TT_HD bool intersect_sphere(float3 org, float3 dir, float4 cr, float tmax, float &t) {
  const double ox = (double)org.x - cr.x, oy = (double)org.y - cr.y, oz = (double)org.z - cr.z;
  const double dx = dir.x, dy = dir.y, dz = dir.z, r = cr.w;
  const double a = dx * dx + dy * dy + dz * dz;
  const double b = ox * dx + oy * dy + oz * dz;
  const double s = b / a;
  const double qx = ox - dx * s, qy = oy - dy * s, qz = oz - dz * s;
  const double disc = a * (r * r - (qx * qx + qy * qy + qz * qz));
  if (disc < 0.) return false;
  const double sd = sqrt(disc);
  const double q = b > 0. ? -b - sd : -b + sd;
  const double c = ox * ox + oy * oy + oz * oz - r * r;
  double t0 = q != 0. ? c / q : 0., t1 = q / a;
  if (t0 > t1) { double x = t0; t0 = t1; t1 = x; }
  const double eps = 1e-9 * ((double)fabsf(cr.x) + fabsf(cr.y) + fabsf(cr.z) + r);
  if (t0 > eps && t0 < (double)tmax) { t = (float)t0; return true; }
  if (t1 > eps && t1 < (double)tmax) { t = (float)t1; return true; }
  return false;
}

TT_HD bool intersect_prim(const PrimRecord &pr, const RayPre &p, const Ray &r, float tmax, float &t, float &u, float &v) {
  if (as_int(pr.b.w) == 0)
    return intersect_triangle(p, r.org, xyz(pr.a), xyz(pr.b), xyz(pr.c), tmax, t, u, v);
  u = v = 0.f;
  return intersect_sphere(r.org, r.dir, make_float4(pr.a.x, pr.a.y, pr.a.z, pr.b.x), tmax, t);
}

// Slab test of both children; returns entry distances (or +inf when missed).
TT_HD void box2(const BvhNode &n, const RayPre &p, float tmax, float &tl, float &tr) {
  float lx0 = n.lxy.x * p.inv.x - p.oinv.x, lx1 = n.lxy.y * p.inv.x - p.oinv.x;
  float ly0 = n.lxy.z * p.inv.y - p.oinv.y, ly1 = n.lxy.w * p.inv.y - p.oinv.y;
  float lz0 = n.lrz.x * p.inv.z - p.oinv.z, lz1 = n.lrz.y * p.inv.z - p.oinv.z;
  float rx0 = n.rxy.x * p.inv.x - p.oinv.x, rx1 = n.rxy.y * p.inv.x - p.oinv.x;
  float ry0 = n.rxy.z * p.inv.y - p.oinv.y, ry1 = n.rxy.w * p.inv.y - p.oinv.y;
  float rz0 = n.lrz.z * p.inv.z - p.oinv.z, rz1 = n.lrz.w * p.inv.z - p.oinv.z;
  // Conservative far bound (1 + 2 gamma(3)) so grazing hits are not lost (Ize 2013).
  const float pad = 1.0000004f;
  float lmin = fmaxf(fmaxf(fminf(lx0, lx1), fminf(ly0, ly1)), fmaxf(fminf(lz0, lz1), 0.f));
  float lmax = fminf(fminf(fmaxf(lx0, lx1), fmaxf(ly0, ly1)), fminf(fmaxf(lz0, lz1), tmax)) * pad;
  float rmin = fmaxf(fmaxf(fminf(rx0, rx1), fminf(ry0, ry1)), fmaxf(fminf(rz0, rz1), 0.f));
  float rmax = fminf(fminf(fmaxf(rx0, rx1), fmaxf(ry0, ry1)), fminf(fmaxf(rz0, rz1), tmax)) * pad;
  tl = lmin <= lmax ? lmin : as_float(0x7f800000);
  tr = rmin <= rmax ? rmin : as_float(0x7f800000);
}

// Generic traversal.  `stack` points at this lane's LDS column (stride kTraceBlock).
template <bool kAnyHit>
TT_HD bool traverse(const BvhNode *__restrict__ nodes, const PrimRecord *__restrict__ prims, const Ray &r,
                                         int *stack, HitRec &hit) {
  RayPre p = precompute(r);
  float tmax = r.tmax;
  hit.prim = -1;
  int spill[kSpillStack];
  int sp = 0;
  int node = 0;
  for (;;) {
    const BvhNode n = nodes[node];
    float tl, tr;
    box2(n, p, tmax, tl, tr);
    int next = -1;
    // Leaves are intersected as soon as their parent is visited.
    int children[2] = {n.ch.x, n.ch.y};
    int counts[2] = {n.ch.z, n.ch.w};
    float tc[2] = {tl, tr};
    int inner[2];
    float tin[2];
    int ninner = 0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (!(tc[c] <= tmax)) continue;
      if (children[c] < 0) {
        int first = ~children[c];
        for (int k = 0; k < counts[c]; ++k) {
          float t, u, v;
          if (intersect_prim(prims[first + k], p, r, tmax, t, u, v)) {
            tmax = t;
            hit.t = t; hit.u = u; hit.v = v; hit.prim = first + k;
            if (kAnyHit) return true;
          }
        }
      } else {
        inner[ninner] = children[c];
        tin[ninner] = tc[c];
        ++ninner;
      }
    }
    if (ninner == 2) {
      bool swap = tin[1] < tin[0];
      next = swap ? inner[1] : inner[0];
      int far = swap ? inner[0] : inner[1];
      // The host rejects trees deeper than kLdsStack + kSpillStack, so this bound never drops a node.
      if (sp < kLdsStack) stack[sp * kTraceBlock] = far;
      else if (sp < kLdsStack + kSpillStack) spill[sp - kLdsStack] = far;
      ++sp;
      if (sp > kLdsStack + kSpillStack) sp = kLdsStack + kSpillStack;
    } else if (ninner == 1) {
      next = inner[0];
    }
    if (next < 0) {
      if (sp == 0) break;
      --sp;
      next = sp < kLdsStack ? stack[sp * kTraceBlock] : spill[sp - kLdsStack];
    }
    node = next;
  }
  return hit.prim >= 0;
}

// Resumable traversal for persistent kernels with dynamic ray fetch: a lane advances its ray one
// node at a time and, when the ray is finished, takes the next one while the rest of the wave keeps
// traversing (persistent while-while with dynamic fetch, Aila & Laine 2009).
struct Trav {
  RayPre p;
  float3 org, dir;
  float tmax;
  HitRec hit;
  int node, sp;
};

TT_HD void trav_init(Trav &t, const Ray &r) {
  t.p = precompute(r);
  t.org = r.org;
  t.dir = r.dir;
  t.tmax = r.tmax;
  t.hit.prim = -1;
  t.node = 0;
  t.sp = 0;
}

// Processes the current node (testing leaf primitives right away); returns false when the
// traversal is finished (stack empty, or an occluder found for any-hit queries).
template <bool kAnyHit>
TT_HD bool trav_step(Trav &t, const BvhNode *__restrict__ nodes, const PrimRecord *__restrict__ prims, int *stack, int *spill) {
  const BvhNode n = nodes[t.node];
  float tl, tr;
  box2(n, t.p, t.tmax, tl, tr);
  const int children[2] = {n.ch.x, n.ch.y};
  const int counts[2] = {n.ch.z, n.ch.w};
  const float tc[2] = {tl, tr};
  int inner[2];
  float tin[2];
  int ninner = 0;
  Ray r;
  r.org = t.org;
  r.dir = t.dir;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (!(tc[c] <= t.tmax)) continue;
    if (children[c] < 0) {
      const int first = ~children[c];
      for (int k = 0; k < counts[c]; ++k) {
        float th, u, v;
        if (intersect_prim(prims[first + k], t.p, r, t.tmax, th, u, v)) {
          t.tmax = th;
          t.hit.t = th; t.hit.u = u; t.hit.v = v; t.hit.prim = first + k;
          if (kAnyHit) return false;
        }
      }
    } else {
      inner[ninner] = children[c];
      tin[ninner] = tc[c];
      ++ninner;
    }
  }
  int next = -1;
  if (ninner == 2) {
    const bool swap = tin[1] < tin[0];
    next = swap ? inner[1] : inner[0];
    const int far = swap ? inner[0] : inner[1];
    // The host rejects trees deeper than kLdsStack + kSpillStack, so this bound never drops a node.
    if (t.sp < kLdsStack) stack[t.sp * kTraceBlock] = far;
    else if (t.sp < kLdsStack + kSpillStack) spill[t.sp - kLdsStack] = far;
    ++t.sp;
    if (t.sp > kLdsStack + kSpillStack) t.sp = kLdsStack + kSpillStack;
  } else if (ninner == 1) {
    next = inner[0];
  }
  if (next < 0) {
    if (t.sp == 0) return false;
    --t.sp;
    next = t.sp < kLdsStack ? stack[t.sp * kTraceBlock] : spill[t.sp - kLdsStack];
  }
  t.node = next;
  return true;
}

}  // namespace ttgpu
