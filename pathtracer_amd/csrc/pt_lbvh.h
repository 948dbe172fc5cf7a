// pt_lbvh.h -- host interface of the on-device linear BVH build (pt_lbvh.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace pt {

// Device scratch for a build over n primitives, carved from one allocation of
// lbvh_scratch_bytes(n) bytes.
struct LbvhScratch {
  float4 *box, *leafbox, *nodebox;   // per-primitive, sorted-primitive and internal-node boxes (2 float4 each)
  uint64_t *key, *key2;              // Morton codes, before and after the sort
  uint32_t *idx, *idx2;              // primitive indices, before and after the sort
  int2 *child, *range;               // binary tree: children and primitive range of each internal node
  int *parent_int, *parent_leaf;     // parent of each internal node / sorted primitive
  uint32_t *flags;                   // refit arrival counters, one per internal node
  int2 *q0, *q1;                     // collapse work lists: (binary node, 4-wide node); q0 first holds the refit frontier
  uint32_t *cb;                      // centroid and scene bounds as ordered uints (12)
  uint32_t *counts;                  // [0] refit frontier length, [1] next work-list length, [2] 4-wide node count
  void *sort_tmp;
  size_t sort_bytes;
};

struct LbvhResult {
  int root;          // 0, or a leaf code when every primitive fits one leaf
  uint32_t nodes4;   // 4-wide nodes written
  uint32_t depth4;   // levels of 4-wide nodes
  uint32_t level_end[4];  // nodes in levels 0..l (nodes are numbered level by level)
  float lo[3], hi[3];  // scene bounds (unpadded)
};

size_t lbvh_scratch_bytes(uint32_t n);
void lbvh_carve(void *base, uint32_t n, LbvhScratch &sc);

// Builds the tree over the n primitive records `prims` (3 float4 each, in id
// order): writes the records in leaf order to prims_out and the 4-wide nodes
// to nodes4, which must hold max(1, n - 1) nodes.  Synchronises `st`.
hipError_t lbvh_build(hipStream_t st, const float4 *prims, uint32_t n, uint32_t leaf_size, LbvhScratch &sc,
                      float4 *prims_out, float4 *nodes4, LbvhResult &res);

// The n 128-byte nodes (8 float4 each) as quantised any-hit traversal nodes
// (PT_QNODE_F4 float4 each, quantize_node).
hipError_t lbvh_quantize(hipStream_t st, const float4 *nodes8, uint32_t n, float4 *nodes_q);

}  // namespace pt
