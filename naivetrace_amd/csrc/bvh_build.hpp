// Host-side binned-SAH BVH builder (replaces Embree's BVH build,
// src/embreeaccelerator.cxx:36-84, and Scene::BuildAccelStructure,
// src/scene.cxx:116-141).  Produces the GPU layout of device_scene.hpp.
#pragma once
#include <cstdint>
#include <vector>

namespace ttgpu {

struct BuildPrim {
  float lo[3], hi[3];
};

struct BuiltBvh {
  // Node array in the device layout (BvhNode, 16 floats each).
  std::vector<float> nodes;      // 16 * num_nodes
  std::vector<uint32_t> order;   // leaf-order position -> input primitive index
  int num_nodes = 0;
  int max_depth = 0;
  float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
};

// SAH with 32 bins per axis, leaves of at most `max_leaf` primitives.
BuiltBvh BuildBvh(const std::vector<BuildPrim> &prims, int max_leaf = 4, int num_threads = 0);

// 4-wide BVH collapsed from a BVH2 (device layout Bvh4Node, 32 words per node): every node takes
// its BVH2 node's two children and then opens the inner child of largest surface area until it has
// four.  Leaves keep the BVH2's primitive ranges (same leaf order).
struct BuiltBvh4 {
  std::vector<float> nodes;      // 32 * num_nodes
  int num_nodes = 0;
  int max_depth = 0;             // depth of the deepest node (root = 0)
};
BuiltBvh4 CollapseToBvh4(const BuiltBvh &b2);

// The same tree with 8-bit child planes (device layout QBvh4Node, 16 words per node): per node and
// axis a grid origin (the lowest child plane) and a power-of-two step chosen so that 255 steps span
// the children; lower planes round down and upper planes up, each checked after the decode the
// device performs, so every decoded child box contains the true one.  Empty slots (leaf, count 0)
// get q = 0.  Throws if a leaf holds more than 255 primitives or an extent exceeds 255 x 2^127.
std::vector<uint32_t> QuantizeBvh4(const BuiltBvh4 &b4);

}  // namespace ttgpu
