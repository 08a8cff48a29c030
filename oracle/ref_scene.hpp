// ORACLE — test infrastructure only (see ref_math.hpp).
//
// Scene model of the oracle: primitives, a median-split BVH in FP64 with a
// watertight ray/triangle test (Woop, Benthin, Wald 2013) and FP64 spheres.
// Replaces EmbreeAccelerator's queries (src/embreeaccelerator.cxx:87-304):
// closest hit, occlusion, and the ordered, de-duplicated list of all hits
// used by the medium tracker (:134-213).
#pragma once
#include "ref_math.hpp"
#include "../include/ttgpu.h"
#include <vector>
#include <memory>

namespace ttref {

struct Prim {
  int kind;        // 0 triangle, 1 sphere
  int index;       // triangle or sphere index in the input
  int material;
};

struct Hit {
  int prim = -1;   // index into Scene::prims
  double t = Infinity;
  double b0 = 0, b1 = 0, b2 = 0;  // barycentrics (triangle); unused for spheres
};

struct BoundaryHit {
  int prim;
  double t;
  V3 n;            // unoriented geometric normal
};

struct Bvh {
  struct Node {
    double lo[3], hi[3];
    int left = -1, right = -1;  // children; leaf when left < 0
    int start = 0, count = 0;   // leaf prim range into order
  };
  std::vector<Node> nodes;
  std::vector<int> order;       // prim ids
};

class Scene {
 public:
  explicit Scene(const ttgpu_scene &s);

  // Closest surface hit in (tnear, tfar).
  bool FirstIntersection(V3 org, V3 dir, double tnear, double tfar, Hit &hit) const;
  bool IsOccluded(V3 org, V3 dir, double tnear, double tfar) const;
  // All hits with volume (or, if `surfaces`, surface) geometry in (tnear, tfar),
  // sorted by t, de-duplicated on equal (material, t) like the reference's filter.
  void IntersectionsInOrder(V3 org, V3 dir, double tnear, double tfar, bool surfaces,
                            std::vector<BoundaryHit> &out) const;

  // Local geometry at a hit: position, geometric/smooth normals, FP32 error bounds.
  void LocalGeometry(const Hit &h, V3 &pos, V3 &gn, V3 &sn, V3 &pos_bounds) const;
  double PrimArea(int prim) const;
  // Uniform position on a primitive (Mesh/Spheres::SampleUniformPosition, primitive.cxx:120-128,193-203).
  Hit SampleUniformPosition(int prim, const double r[2]) const;

  std::vector<Prim> prims;
  std::vector<float> vtx, nrm;
  std::vector<uint32_t> idx;
  std::vector<float> sph;
  std::vector<int> surfaces, volumes, area_lights;
  std::vector<int> area_light_of_prim;  // -1 if not emissive
  Bvh bvh_surf, bvh_vol;
  double bb_lo[3], bb_hi[3];
  bool has_bounds = false;

  // Copies of the tables.
  std::vector<ttgpu_shader> shaders;
  std::vector<ttgpu_medium> media;
  std::vector<ttgpu_material> materials;
  std::vector<ttgpu_area_emitter> emitters;
  std::vector<ttgpu_env_light> envs;
  std::vector<std::vector<float>> env_textures;
  std::vector<ttgpu_point_light> points;
  ttgpu_camera camera;
  int empty_space_medium = -1;

  V3 TriVertex(int tri, int k) const {
    uint32_t i = idx[3 * tri + k];
    return {vtx[3 * i], vtx[3 * i + 1], vtx[3 * i + 2]};
  }
  double BoundingBoxMaxExtent() const;
  double BoundingBoxDiagonal() const;
  bool InBox(V3 p) const;

 private:
  void Build(Bvh &b, const std::vector<int> &ids);
  int BuildRec(Bvh &b, std::vector<int> &ids, int begin, int end, std::vector<double> &centroids, std::vector<double> &boxes);
  bool IntersectPrim(int pid, V3 org, V3 dir, double tnear, double tfar, Hit &h) const;
  template <class F> void Traverse(const Bvh &b, V3 org, V3 dir, double tnear, double &tfar, F &&leaf) const;
};

}  // namespace ttref
