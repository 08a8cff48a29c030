// ORACLE — test infrastructure only (see ref_math.hpp).
#include "ref_scene.hpp"
#include <stdexcept>
#include <numeric>

namespace ttref {

Scene::Scene(const ttgpu_scene &s) {
  if (s.num_triangles > 0 && (!s.vertices || !s.indices || !s.tri_material))
    throw std::invalid_argument("triangles without vertex/index/material arrays");
  vtx.assign(s.vertices, s.vertices + 3 * s.num_vertices);
  if (s.normals) nrm.assign(s.normals, s.normals + 3 * s.num_vertices);
  idx.assign(s.indices, s.indices + 3 * s.num_triangles);
  if (s.num_spheres > 0) sph.assign(s.spheres, s.spheres + 4 * s.num_spheres);
  shaders.assign(s.shaders, s.shaders + s.num_shaders);
  media.assign(s.media, s.media + s.num_media);
  materials.assign(s.materials, s.materials + s.num_materials);
  emitters.assign(s.area_emitters, s.area_emitters + s.num_area_emitters);
  envs.assign(s.env_lights, s.env_lights + s.num_env_lights);
  for (auto &e : envs) {
    if (e.kind == TTGPU_ENV_MAP) {
      env_textures.emplace_back(e.texture_rgb, e.texture_rgb + 3 * (size_t)e.tex_width * e.tex_height);
      e.texture_rgb = env_textures.back().data();
    }
  }
  points.assign(s.point_lights, s.point_lights + s.num_point_lights);
  camera = s.camera;
  empty_space_medium = s.empty_space_medium;

  for (int64_t i = 0; i < s.num_triangles; ++i) {
    for (int k = 0; k < 3; ++k)
      if (idx[3 * i + k] >= (uint32_t)s.num_vertices) throw std::invalid_argument("triangle index out of range");
    prims.push_back({0, (int)i, s.tri_material[i]});
  }
  for (int64_t i = 0; i < s.num_spheres; ++i) prims.push_back({1, (int)i, s.sphere_material[i]});
  area_light_of_prim.assign(prims.size(), -1);
  for (size_t p = 0; p < prims.size(); ++p) {
    int m = prims[p].material;
    if (m < 0 || m >= (int)materials.size()) throw std::invalid_argument("material index out of range");
    const auto &mat = materials[m];
    if (mat.shader >= 0) {
      surfaces.push_back((int)p);
      if (mat.emitter >= 0) {
        area_light_of_prim[p] = (int)area_lights.size();
        area_lights.push_back((int)p);
      }
    } else {
      if (mat.medium < 0 && mat.outer_medium < 0) throw std::invalid_argument("material without shader and medium");
      volumes.push_back((int)p);
    }
  }
  Build(bvh_surf, surfaces);
  Build(bvh_vol, volumes);
  // Scene bounds over surfaces and volumes (Scene::BuildAccelStructure, scene.cxx:139-140).
  for (int k = 0; k < 3; ++k) { bb_lo[k] = Infinity; bb_hi[k] = -Infinity; }
  for (const Bvh *b : {&bvh_surf, &bvh_vol}) {
    if (b->nodes.empty()) continue;
    has_bounds = true;
    for (int k = 0; k < 3; ++k) {
      bb_lo[k] = std::min(bb_lo[k], b->nodes[0].lo[k]);
      bb_hi[k] = std::max(bb_hi[k], b->nodes[0].hi[k]);
    }
  }
}

double Scene::BoundingBoxMaxExtent() const {
  return std::max({bb_hi[0] - bb_lo[0], bb_hi[1] - bb_lo[1], bb_hi[2] - bb_lo[2]});
}
double Scene::BoundingBoxDiagonal() const {
  return std::sqrt(Sqr(bb_hi[0] - bb_lo[0]) + Sqr(bb_hi[1] - bb_lo[1]) + Sqr(bb_hi[2] - bb_lo[2]));
}
bool Scene::InBox(V3 p) const {
  if (!has_bounds) return false;
  for (int k = 0; k < 3; ++k)
    if (p[k] < bb_lo[k] || p[k] > bb_hi[k]) return false;
  return true;
}

static void PrimBox(const Scene &sc, int pid, double lo[3], double hi[3]) {
  const Prim &p = sc.prims[pid];
  if (p.kind == 0) {
    for (int k = 0; k < 3; ++k) { lo[k] = Infinity; hi[k] = -Infinity; }
    for (int j = 0; j < 3; ++j) {
      V3 v = sc.TriVertex(p.index, j);
      for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], v[k]); hi[k] = std::max(hi[k], v[k]); }
    }
  } else {
    const float *s = &sc.sph[4 * p.index];
    for (int k = 0; k < 3; ++k) { lo[k] = (double)s[k] - s[3]; hi[k] = (double)s[k] + s[3]; }
  }
}

void Scene::Build(Bvh &b, const std::vector<int> &ids) {
  b.nodes.clear();
  b.order = ids;
  if (ids.empty()) return;
  std::vector<double> centroids(3 * prims.size()), boxes(6 * prims.size());
  for (int pid : ids) {
    double lo[3], hi[3];
    PrimBox(*this, pid, lo, hi);
    for (int k = 0; k < 3; ++k) {
      centroids[3 * pid + k] = 0.5 * (lo[k] + hi[k]);
      boxes[6 * pid + k] = lo[k];
      boxes[6 * pid + 3 + k] = hi[k];
    }
  }
  b.nodes.reserve(2 * ids.size());
  BuildRec(b, b.order, 0, (int)ids.size(), centroids, boxes);
}

// Object-median split on the longest centroid axis, leaves of <= 4 primitives.
int Scene::BuildRec(Bvh &b, std::vector<int> &ids, int begin, int end, std::vector<double> &cen, std::vector<double> &box) {
  int ni = (int)b.nodes.size();
  b.nodes.emplace_back();
  Bvh::Node nd;
  double clo[3] = {Infinity, Infinity, Infinity}, chi[3] = {-Infinity, -Infinity, -Infinity};
  for (int k = 0; k < 3; ++k) { nd.lo[k] = Infinity; nd.hi[k] = -Infinity; }
  for (int i = begin; i < end; ++i) {
    int p = ids[i];
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] = std::min(nd.lo[k], box[6 * p + k]);
      nd.hi[k] = std::max(nd.hi[k], box[6 * p + 3 + k]);
      clo[k] = std::min(clo[k], cen[3 * p + k]);
      chi[k] = std::max(chi[k], cen[3 * p + k]);
    }
  }
  int n = end - begin;
  if (n <= 4) {
    nd.start = begin;
    nd.count = n;
    b.nodes[ni] = nd;
    return ni;
  }
  int axis = 0;
  for (int k = 1; k < 3; ++k)
    if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
  int mid = begin + n / 2;
  std::nth_element(ids.begin() + begin, ids.begin() + mid, ids.begin() + end,
                   [&](int a, int c) { return cen[3 * a + axis] < cen[3 * c + axis]; });
  int l = BuildRec(b, ids, begin, mid, cen, box);
  int r = BuildRec(b, ids, mid, end, cen, box);
  nd.left = l;
  nd.right = r;
  b.nodes[ni] = nd;
  return ni;
}

static bool RayBox(const double lo[3], const double hi[3], V3 org, V3 inv, double tnear, double tfar, double &tent) {
  double t0 = tnear, t1 = tfar;
  for (int k = 0; k < 3; ++k) {
    double a = (lo[k] - org[k]) * inv[k], c = (hi[k] - org[k]) * inv[k];
    if (a > c) std::swap(a, c);
    // Robust slab bounds: pad the far distance (Ize 2013).
    c *= 1 + 2 * GammaF(3);
    if (std::isnan(a)) a = -Infinity;
    if (std::isnan(c)) c = Infinity;
    t0 = std::max(t0, a);
    t1 = std::min(t1, c);
  }
  tent = t0;
  return t0 <= t1;
}

template <class F> void Scene::Traverse(const Bvh &b, V3 org, V3 dir, double tnear, double &tfar, F &&leaf) const {
  if (b.nodes.empty()) return;
  V3 inv{1. / dir.x, 1. / dir.y, 1. / dir.z};
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const Bvh::Node &nd = b.nodes[stack[--sp]];
    double te;
    if (!RayBox(nd.lo, nd.hi, org, inv, tnear, tfar, te)) continue;
    if (nd.left < 0) {
      for (int i = nd.start; i < nd.start + nd.count; ++i)
        if (leaf(b.order[i])) return;
    } else {
      stack[sp++] = nd.right;
      stack[sp++] = nd.left;
    }
  }
}

// Watertight ray/triangle intersection, Woop, Benthin & Wald (2013), in FP64.
static bool TriWatertight(V3 org, V3 dir, V3 A, V3 B, V3 C, double tnear, double tfar, double &t, double &b0, double &b1, double &b2) {
  int kz = 0;
  if (std::abs(dir.y) > std::abs(dir[kz])) kz = 1;
  if (std::abs(dir.z) > std::abs(dir[kz])) kz = 2;
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  if (dir[kz] < 0) std::swap(kx, ky);
  const double Sx = dir[kx] / dir[kz], Sy = dir[ky] / dir[kz], Sz = 1. / dir[kz];
  V3 a = A - org, bb = B - org, c = C - org;
  const double Ax = a[kx] - Sx * a[kz], Ay = a[ky] - Sy * a[kz];
  const double Bx = bb[kx] - Sx * bb[kz], By = bb[ky] - Sy * bb[kz];
  const double Cx = c[kx] - Sx * c[kz], Cy = c[ky] - Sy * c[kz];
  double U = Cx * By - Cy * Bx;
  double V = Ax * Cy - Ay * Cx;
  double W = Bx * Ay - By * Ax;
  if (U == 0. || V == 0. || W == 0.) {
    long double CxBy = (long double)Cx * By, CyBx = (long double)Cy * Bx;
    U = (double)(CxBy - CyBx);
    long double AxCy = (long double)Ax * Cy, AyCx = (long double)Ay * Cx;
    V = (double)(AxCy - AyCx);
    long double BxAy = (long double)Bx * Ay, ByAx = (long double)By * Ax;
    W = (double)(BxAy - ByAx);
  }
  if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return false;
  const double det = U + V + W;
  if (det == 0.) return false;
  const double Az = Sz * a[kz], Bz = Sz * bb[kz], Cz = Sz * c[kz];
  const double T = U * Az + V * Bz + W * Cz;
  const double inv = 1. / det;
  t = T * inv;
  if (!(t > tnear && t < tfar)) return false;
  // Conservative lower bound on t (PBRT-v3, Triangle::Intersect): a hit closer than the rounding
  // error of t is rejected, so rays leaving a surface do not re-hit it at t ~ 0.
  {
    const double g = std::numeric_limits<double>::epsilon();
    auto gam = [g](int n) { return n * g / (1 - n * g); };
    const double maxZt = std::max({std::abs(Az), std::abs(Bz), std::abs(Cz)});
    const double maxXt = std::max({std::abs(Ax), std::abs(Bx), std::abs(Cx)});
    const double maxYt = std::max({std::abs(Ay), std::abs(By), std::abs(Cy)});
    const double deltaZ = gam(3) * maxZt, deltaX = gam(5) * (maxXt + maxZt), deltaY = gam(5) * (maxYt + maxZt);
    const double deltaE = 2 * (gam(2) * maxXt * maxYt + deltaY * maxXt + deltaX * maxYt);
    const double maxE = std::max({std::abs(U), std::abs(V), std::abs(W)});
    const double deltaT = 3 * (gam(3) * maxE * maxZt + deltaE * maxZt + deltaZ * maxE) * std::abs(inv);
    if (t <= deltaT) return false;
  }
  // Barycentrics relative to vertices (A,B,C): p = b0*A + b1*B + b2*C.
  b0 = U * inv;
  b1 = V * inv;
  b2 = W * inv;
  return true;
}

bool Scene::IntersectPrim(int pid, V3 org, V3 dir, double tnear, double tfar, Hit &h) const {
  const Prim &p = prims[pid];
  if (p.kind == 0) {
    double t, b0, b1, b2;
    if (!TriWatertight(org, dir, TriVertex(p.index, 0), TriVertex(p.index, 1), TriVertex(p.index, 2), tnear, tfar, t, b0, b1, b2))
      return false;
    h.prim = pid; h.t = t; h.b0 = b0; h.b1 = b1; h.b2 = b2;
    return true;
  }
  const float *s = &sph[4 * p.index];
  V3 ctr{s[0], s[1], s[2]};
  V3 v = org - ctr;
  double A = Dot(dir, dir), B = 2. * Dot(v, dir), C = Dot(v, v) - Sqr((double)s[3]);
  double t0, t1;
  if (!Quadratic(A, B, C, t0, t1)) return false;
  double t = (t0 > tnear && t0 < tfar) ? t0 : ((t1 > tnear && t1 < tfar) ? t1 : -1.);
  if (t < 0.) return false;
  V3 off = Normalized(org + dir * t - ctr) * (double)s[3];
  h.prim = pid; h.t = t; h.b0 = off.x; h.b1 = off.y; h.b2 = off.z;
  return true;
}

bool Scene::FirstIntersection(V3 org, V3 dir, double tnear, double tfar, Hit &hit) const {
  bool found = false;
  double tf = tfar;
  Traverse(bvh_surf, org, dir, tnear, tf, [&](int pid) {
    Hit h;
    if (IntersectPrim(pid, org, dir, tnear, tf, h)) {
      if (h.t < tf || !found) { hit = h; tf = h.t; found = true; }
    }
    return false;
  });
  return found;
}

bool Scene::IsOccluded(V3 org, V3 dir, double tnear, double tfar) const {
  bool occ = false;
  double tf = tfar;
  Traverse(bvh_surf, org, dir, tnear, tf, [&](int pid) {
    Hit h;
    if (IntersectPrim(pid, org, dir, tnear, tfar, h)) { occ = true; return true; }
    return false;
  });
  return occ;
}

void Scene::IntersectionsInOrder(V3 org, V3 dir, double tnear, double tfar, bool surf, std::vector<BoundaryHit> &out) const {
  out.clear();
  double tf = tfar;
  const Bvh &b = surf ? bvh_surf : bvh_vol;
  auto add = [&](int pid, double t) {
    V3 pos, gn, sn, pb;
    Hit h;
    h.prim = pid;
    h.t = t;
    // Normal only needed: compute from primitive.
    const Prim &p = prims[pid];
    if (p.kind == 0) {
      V3 A = TriVertex(p.index, 0), B = TriVertex(p.index, 1), C = TriVertex(p.index, 2);
      gn = Normalized(Cross(B - A, C - A));
    } else {
      const float *s = &sph[4 * p.index];
      gn = Normalized(org + dir * t - V3{s[0], s[1], s[2]});
    }
    BoundaryHit bh{pid, t, gn};
    auto it = std::lower_bound(out.begin(), out.end(), bh, [](const BoundaryHit &a, const BoundaryHit &c) { return a.t < c.t; });
    for (auto jt = it; jt != out.end() && jt->t == t; ++jt)
      if (prims[jt->prim].material == p.material) return;
    out.insert(it, bh);
  };
  Traverse(b, org, dir, tnear, tf, [&](int pid) {
    const Prim &p = prims[pid];
    if (p.kind == 0) {
      Hit h;
      if (IntersectPrim(pid, org, dir, tnear, tfar, h)) add(pid, h.t);
    } else {
      const float *s = &sph[4 * p.index];
      V3 v = org - V3{s[0], s[1], s[2]};
      double A = Dot(dir, dir), B = 2. * Dot(v, dir), C = Dot(v, v) - Sqr((double)s[3]);
      double t0, t1;
      if (Quadratic(A, B, C, t0, t1)) {
        if (t0 > tnear && t0 < tfar) add(pid, t0);
        if (t1 > tnear && t1 < tfar) add(pid, t1);
      }
    }
    return false;
  });
}

void Scene::LocalGeometry(const Hit &h, V3 &pos, V3 &gn, V3 &sn, V3 &pb) const {
  const Prim &p = prims[h.prim];
  if (p.kind == 0) {
    V3 A = TriVertex(p.index, 0), B = TriVertex(p.index, 1), C = TriVertex(p.index, 2);
    pos = A * h.b0 + B * h.b1 + C * h.b2;
    gn = Normalized(Cross(B - A, C - A));
    if (!nrm.empty()) {
      V3 n[3];
      for (int k = 0; k < 3; ++k) {
        uint32_t i = idx[3 * p.index + k];
        n[k] = V3{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]};
      }
      sn = Normalized(n[0] * h.b0 + n[1] * h.b1 + n[2] * h.b2);
    } else {
      sn = gn;
    }
    // Error bounds, PBRT 3.9 (FillPosBoundsTriangle, primitive.cxx:291-299).
    const double g = GammaF(7);
    for (int k = 0; k < 3; ++k) pb[k] = (std::abs(A[k] * h.b0) + std::abs(B[k] * h.b1) + std::abs(C[k] * h.b2)) * g;
  } else {
    // Spheres: the "barycentrics" hold the offset from the centre (Spheres::GetLocalGeometry,
    // primitive.cxx:214-223); error bounds PBRT pg. 225 (FillPosBoundsSphere).
    const float *s = &sph[4 * p.index];
    V3 off{h.b0, h.b1, h.b2};
    pos = V3{s[0], s[1], s[2]} + off;
    gn = sn = Normalized(off);
    const double g = GammaF(5);
    for (int k = 0; k < 3; ++k) pb[k] = std::abs(pos[k]) * g;
  }
}

double Scene::PrimArea(int pid) const {
  const Prim &p = prims[pid];
  if (p.kind == 0) {
    // Mesh::Area computes the cross product in float (primitive.cxx:130-138).
    float a[3], b[3], c[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = vtx[3 * idx[3 * p.index] + k];
      b[k] = vtx[3 * idx[3 * p.index + 1] + k];
      c[k] = vtx[3 * idx[3 * p.index + 2] + k];
    }
    float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    return 0.5 * std::sqrt((double)n[0] * n[0] + (double)n[1] * n[1] + (double)n[2] * n[2]);
  }
  double r = sph[4 * p.index + 3];
  return Sqr(r) * UnitSphereSurfaceArea;
}

Hit Scene::SampleUniformPosition(int pid, const double r[2]) const {
  Hit h;
  h.prim = pid;
  const Prim &p = prims[pid];
  if (p.kind == 0) {
    V3 b = ToTriangleBarycentricCoords(r);
    h.b0 = b.x; h.b1 = b.y; h.b2 = b.z;
  } else {
    // For spheres the "barycentrics" carry the offset from the centre (primitive.cxx:193-203).
    V3 d = ToUniformSphere(r) * (double)sph[4 * p.index + 3];
    h.b0 = d.x; h.b1 = d.y; h.b2 = d.z;
  }
  return h;
}

}  // namespace ttref
