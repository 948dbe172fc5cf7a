// pt_lbvh.hip -- on-device BVH build (linear BVH, Karras 2012) for scenes
// whose instances change often: primitive bounds -> 63-bit Morton codes of
// the centroids -> radix sort -> the binary radix tree of the sorted codes
// (one thread per internal node) -> bottom-up box refit -> collapse into the
// renderer's 4-wide nodes, level by level, with subtrees of at most leaf_size
// primitives as leaves.  The primitives are reordered into leaf order on the
// device.  The result has the same layout as the host SAH build (pt_scene.h),
// so traversal is unchanged.
#include <hipcub/hipcub.hpp>

#include "pt_device.h"
#include "pt_lbvh.h"

namespace pt {

__device__ __forceinline__ uint32_t l_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Order-preserving float <-> uint for atomic min / max.
PT_D uint32_t f2o(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
PT_D float o2f(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

// Bounds of primitive record i (the prims layout of pt_scene.h), before padding.
PT_D void prim_bounds(const float4 *P, f3 &lo, f3 &hi) {
  const float4 A = P[0], B = P[1], C = P[2];
  const uint32_t type = __float_as_uint(A.w) >> 30;
  if (type == PRIM_TRI) {
    const f3 p0 = mk3(A.x, A.y, A.z), p1 = add3(p0, mk3(B.x, B.y, B.z)), p2 = add3(p0, mk3(C.x, C.y, C.z));
    lo = mk3(fminf(p0.x, fminf(p1.x, p2.x)), fminf(p0.y, fminf(p1.y, p2.y)), fminf(p0.z, fminf(p1.z, p2.z)));
    hi = mk3(fmaxf(p0.x, fmaxf(p1.x, p2.x)), fmaxf(p0.y, fmaxf(p1.y, p2.y)), fmaxf(p0.z, fmaxf(p1.z, p2.z)));
  } else if (type == PRIM_SPHERE) {
    const f3 c = mk3(A.x, A.y, A.z);
    lo = sub3(c, s3(B.x));
    hi = add3(c, s3(B.x));
  } else if (type == PRIM_RECT) {
    const int ax = (int)B.z, b0 = ax == 0 ? 1 : 0, b1 = ax == 2 ? 1 : 2;
    float l[3], h[3];
    l[ax] = h[ax] = A.x;
    l[b0] = A.y; h[b0] = A.z;
    l[b1] = B.x; h[b1] = B.y;
    lo = mk3(l[0], l[1], l[2]);
    hi = mk3(h[0], h[1], h[2]);
  } else {  // curve segment
    const f3 pa = mk3(A.x, A.y, A.z), pb = add3(pa, scl3(mk3(B.x, B.y, B.z), C.x));
    lo = sub3(mk3(fminf(pa.x, pb.x), fminf(pa.y, pb.y), fminf(pa.z, pb.z)), s3(B.w));
    hi = add3(mk3(fmaxf(pa.x, pb.x), fmaxf(pa.y, pb.y), fmaxf(pa.z, pb.z)), s3(B.w));
  }
}

// The same relative padding as the host build: boxes stay conservative under
// the slab test's approximate reciprocal.
PT_D void pad_box(f3 &lo, f3 &hi) {
  const float px = 1e-6f * (fabsf(lo.x) + fabsf(hi.x)) + 1e-12f;
  const float py = 1e-6f * (fabsf(lo.y) + fabsf(hi.y)) + 1e-12f;
  const float pz = 1e-6f * (fabsf(lo.z) + fabsf(hi.z)) + 1e-12f;
  lo = mk3(lo.x - px, lo.y - py, lo.z - pz);
  hi = mk3(hi.x + px, hi.y + py, hi.z + pz);
}

// Per-primitive boxes, and the centroid and scene bounds: a block reduces its
// 256 primitives (wave shuffles, then LDS) before one atomic per bound, so
// the twelve counters see one atomic per block, not per primitive.
// cb = [centroid min xyz, centroid max xyz, scene min xyz, scene max xyz].
__global__ __launch_bounds__(256) void k_lb_bounds(const float4 *prims, uint32_t n, float4 *box, uint32_t *cb) {
  const uint32_t i = l_tid();
  uint32_t v[12];
  for (int k = 0; k < 12; ++k) v[k] = ((k % 6) < 3) ? 0xffffffffu : 0u;  // identities of min / max
  if (i < n) {
    f3 lo, hi;
    prim_bounds(prims + 3 * (size_t)i, lo, hi);
    box[2 * (size_t)i] = make_float4(lo.x, lo.y, lo.z, 0.f);
    box[2 * (size_t)i + 1] = make_float4(hi.x, hi.y, hi.z, 0.f);
    const f3 c = scl3(add3(lo, hi), 0.5f);
    v[0] = v[3] = f2o(c.x);
    v[1] = v[4] = f2o(c.y);
    v[2] = v[5] = f2o(c.z);
    v[6] = f2o(lo.x); v[7] = f2o(lo.y); v[8] = f2o(lo.z);
    v[9] = f2o(hi.x); v[10] = f2o(hi.y); v[11] = f2o(hi.z);
  }
  for (int off = 32; off >= 1; off >>= 1)
    for (int k = 0; k < 12; ++k) {
      const uint32_t o = (uint32_t)__shfl_xor((int)v[k], off);
      v[k] = ((k % 6) < 3) ? min(v[k], o) : max(v[k], o);
    }
  __shared__ uint32_t red[4][12];
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0)
    for (int k = 0; k < 12; ++k) red[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 12) {
    const int k = threadIdx.x;
    uint32_t r = red[0][k];
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) r = ((k % 6) < 3) ? min(r, red[w][k]) : max(r, red[w][k]);
    if ((k % 6) < 3) atomicMin(&cb[k], r); else atomicMax(&cb[k], r);
  }
}

__global__ void k_lb_init_bounds(uint32_t *cb) {
  const uint32_t i = threadIdx.x;
  if (i < 12) cb[i] = ((i % 6) < 3) ? 0xffffffffu : 0u;
}

PT_D uint64_t spread21(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// 63-bit Morton code of the centroid in the centroid bounds (21 bits per
// axis: a small object next to a large one still gets distinct codes).
__global__ void k_lb_morton(const float4 *box, uint32_t n, const uint32_t *cb, uint64_t *code, uint32_t *idx) {
  const uint32_t i = l_tid();
  if (i >= n) return;
  const float4 lo = box[2 * (size_t)i], hi = box[2 * (size_t)i + 1];
  const f3 c = mk3(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z));
  const f3 mn = mk3(o2f(cb[0]), o2f(cb[1]), o2f(cb[2])), mx = mk3(o2f(cb[3]), o2f(cb[4]), o2f(cb[5]));
  auto q = [](float v, float a, float b) {
    const float e = b - a;
    const float t = e > 0.f ? (v - a) / e : 0.5f;
    return (uint64_t)fminf(fmaxf(t * 2097152.f, 0.f), 2097151.f);
  };
  code[i] = (spread21(q(c.x, mn.x, mx.x)) << 2) | (spread21(q(c.y, mn.y, mx.y)) << 1) | spread21(q(c.z, mn.z, mx.z));
  idx[i] = i;
}

// delta of Karras 2012: common prefix length of sorted keys i and j, with the
// index as a tie-break for equal keys; -1 outside the array.
PT_D int delta(const uint64_t *k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = k[i], b = k[j];
  if (a == b) return 64 + __clz((uint32_t)i ^ (uint32_t)j);
  return __clzll(a ^ b);
}

// Internal node i of n - 1: its range and children.  Child codes: >= 0
// internal node, < 0 leaf ~k (sorted primitive k).
__global__ void k_lb_karras(const uint64_t *key, int n, int2 *child, int2 *range, int *parent_int,
                            int *parent_leaf) {
  const int i = (int)l_tid();
  if (i >= n - 1) return;
  const int d = (delta(key, n, i, i + 1) - delta(key, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(key, n, i, i - d);
  int lmax = 2;
  while (delta(key, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(key, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(key, n, i, j);
  int s = 0;
  for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
    if (delta(key, n, i, i + (s + t) * d) > dnode) s += t;
    if (t == 1) break;
  }
  const int g = i + s * d + (d < 0 ? -1 : 0);
  const int lo = min(i, j), hi = max(i, j);
  const int c0 = (lo == g) ? ~g : g;
  const int c1 = (hi == g + 1) ? ~(g + 1) : g + 1;
  child[i] = make_int2(c0, c1);
  range[i] = make_int2(lo, hi);
  if (c0 < 0) parent_leaf[~c0] = i; else parent_int[c0] = i;
  if (c1 < 0) parent_leaf[~c1] = i; else parent_int[c1] = i;
}

// Boxes of the internal nodes, bottom-up: the second child to arrive at a
// node computes its box and goes on.  The same pass takes the leaf decisions
// that minimise the surface-area cost of each subtree (traversal and
// primitive test both cost 1, as in the host build): a subtree of at most
// leaf_size primitives becomes one leaf when that costs no more than its best
// split.  lo.w of a node's box holds the subtree's cost, hi.w 1 for a leaf.
PT_D void child_box(int c, const float4 *leafbox, const float4 *nodebox, float4 &lo, float4 &hi) {
  const float4 *b = c < 0 ? leafbox + 2 * (size_t)(~c) : nodebox + 2 * (size_t)c;
  lo = b[0];
  hi = b[1];
}
PT_D float half_area(const float4 &lo, const float4 &hi) {
  const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
  return dx * dy + dy * dz + dz * dx;
}
PT_D void refit_node(int p, const int2 *child, const int2 *range, const float4 *leafbox, float4 *nodebox,
                     uint32_t leaf_size) {
  const int2 c = child[p];
  float4 l0, h0, l1, h1;
  child_box(c.x, leafbox, nodebox, l0, h0);
  child_box(c.y, leafbox, nodebox, l1, h1);
  const float4 lo = make_float4(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z), 0.f);
  const float4 hi = make_float4(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z), 0.f);
  const float c0 = c.x < 0 ? 1.f : l0.w, c1 = c.y < 0 ? 1.f : l1.w;
  const float ap = half_area(lo, hi);
  const float split = 1.f + (ap > 0.f ? (half_area(l0, h0) * c0 + half_area(l1, h1) * c1) / ap : c0 + c1);
  const int2 r = range[p];
  const int size = r.y - r.x + 1;
  const bool leaf = size <= (int)leaf_size && (float)size <= split;
  nodebox[2 * (size_t)p] = make_float4(lo.x, lo.y, lo.z, leaf ? (float)size : split);
  nodebox[2 * (size_t)p + 1] = make_float4(hi.x, hi.y, hi.z, leaf ? 1.f : 0.f);
}

// Phase 1: the nodes whose primitives all belong to one block's 256 leaves,
// synchronised inside the block (workgroup-scope fences, no cache flushes).
// A thread whose next node reaches beyond its block's leaves puts its
// finished subtree on the frontier list instead.
__global__ __launch_bounds__(256) void k_lb_refit_local(int n, const int2 *child, const int2 *range,
                                                        const int *parent_int, const int *parent_leaf,
                                                        const float4 *leafbox, float4 *nodebox, uint32_t *flags,
                                                        uint32_t leaf_size, int *frontier, uint32_t *n_frontier) {
  const int k = (int)l_tid();
  if (k >= n) return;
  const int b0 = (int)(blockIdx.x * blockDim.x), b1 = min(n, b0 + (int)blockDim.x) - 1;
  int done = ~k;
  int p = parent_leaf[k];
  // a radix tree over 63-bit codes with index tie-breaks over n < 2^27
  // primitives is at most 90 deep
  for (int level = 0; level < 96 && p >= 0 && p < n - 1; ++level) {
    const int2 r = range[p];
    if (r.x < b0 || r.y > b1) {
      frontier[atomicAdd(n_frontier, 1u)] = done;
      return;
    }
    __threadfence_block();
    if (atomicAdd(&flags[p], 1u) == 0u) return;  // the other child is not done yet
    __threadfence_block();
    refit_node(p, child, range, leafbox, nodebox, leaf_size);
    done = p;
    p = p == 0 ? -1 : parent_int[p];
  }
}

// Phase 2: from the frontier up, across blocks: the second child to arrive
// at a node computes it and goes on (device-scope fences around an atomic
// counter per node).
__global__ __launch_bounds__(256) void k_lb_refit_global(int n, const int *frontier, uint32_t n_frontier,
                                                         const int2 *child, const int2 *range,
                                                         const int *parent_int, const int *parent_leaf,
                                                         const float4 *leafbox, float4 *nodebox, uint32_t *flags,
                                                         uint32_t leaf_size) {
  const uint32_t i = l_tid();
  if (i >= n_frontier) return;
  const int f = frontier[i];
  int p = f < 0 ? parent_leaf[~f] : (f == 0 ? -1 : parent_int[f]);
  for (int level = 0; level < 96 && p >= 0 && p < n - 1; ++level) {
    __threadfence();
    if (atomicAdd(&flags[p], 1u) == 0u) return;
    __threadfence();
    refit_node(p, child, range, leafbox, nodebox, leaf_size);
    p = p == 0 ? -1 : parent_int[p];
  }
}

// Leaf order: primitive records and their boxes by sorted index.
__global__ void k_lb_reorder(const float4 *prims, const float4 *box, const uint32_t *idx, uint32_t n,
                             float4 *prims_out, float4 *leafbox) {
  const uint32_t k = l_tid();
  if (k >= n) return;
  const uint32_t i = idx[k];
  for (int q = 0; q < 3; ++q) prims_out[3 * (size_t)k + q] = prims[3 * (size_t)i + q];
  leafbox[2 * (size_t)k] = box[2 * (size_t)i];
  leafbox[2 * (size_t)k + 1] = box[2 * (size_t)i + 1];
}

// One level of the collapse into 4-wide nodes.  Work item w: binary internal
// node q[w].x becomes 4-wide node q[w].y.  Its slots start as the node's two
// children; while fewer than four, the largest-area child that is not a leaf
// is replaced by its own two children (the host collapse's rule).  Children
// that stay internal become the next level's work items.
__global__ void k_lb_collapse(const int2 *qin, uint32_t nq, const int2 *child, const int2 *range,
                              const float4 *leafbox, const float4 *nodebox, float4 *nodes4, uint32_t *n4,
                              int2 *qout, uint32_t *nqout) {
  const uint32_t w = l_tid();
  if (w >= nq) return;
  const int2 it = qin[w];
  auto is_leaf = [&](int x) { return x < 0 || nodebox[2 * (size_t)x + 1].w != 0.f; };
  int items[4];
  const int2 c = child[it.x];
  items[0] = c.x;
  items[1] = c.y;
  int ni = 2;
  while (ni < 4) {
    int best = -1;
    float ba = -1.f;
    for (int k = 0; k < ni; ++k) {
      if (is_leaf(items[k])) continue;
      const float a = half_area(nodebox[2 * (size_t)items[k]], nodebox[2 * (size_t)items[k] + 1]);
      if (a > ba) {
        ba = a;
        best = k;
      }
    }
    if (best < 0) break;
    const int2 g = child[items[best]];
    items[best] = g.x;
    items[ni++] = g.y;
  }
  float lo[3][4], hi[3][4];
  int codes[4];
  for (int k = 0; k < 4; ++k) {
    if (k >= ni) {  // empty slot: an inverted box, as the host build writes
      codes[k] = kEmptyChild;
      for (int a = 0; a < 3; ++a) {
        lo[a][k] = 1.f;
        hi[a][k] = -1.f;
      }
      continue;
    }
    const int x = items[k];
    float4 l4, h4;
    child_box(x, leafbox, nodebox, l4, h4);
    f3 l = mk3(l4.x, l4.y, l4.z), h = mk3(h4.x, h4.y, h4.z);
    pad_box(l, h);
    lo[0][k] = l.x; lo[1][k] = l.y; lo[2][k] = l.z;
    hi[0][k] = h.x; hi[1][k] = h.y; hi[2][k] = h.z;
    if (is_leaf(x)) {
      const int first = x < 0 ? ~x : range[x].x;
      const int sz = x < 0 ? 1 : range[x].y - range[x].x + 1;
      codes[k] = leaf_code((uint32_t)first, (uint32_t)sz);
    } else {
      const uint32_t j = atomicAdd(n4, 1u);
      codes[k] = (int)j;
      qout[atomicAdd(nqout, 1u)] = make_int2(x, (int)j);
    }
  }
  float4 *nd = nodes4 + 8 * (size_t)it.y;
  nd[0] = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
  nd[1] = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
  nd[2] = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
  nd[3] = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
  nd[4] = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
  nd[5] = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
  nd[6] = make_float4(__int_as_float(codes[0]), __int_as_float(codes[1]), __int_as_float(codes[2]),
                      __int_as_float(codes[3]));
  nd[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_lb_root(int2 *q, uint32_t *nq, uint32_t *n4) {
  if (threadIdx.x == 0) {
    q[0] = make_int2(0, 0);
    *nq = 1u;
    *n4 = 1u;
  }
}

// A single primitive: node 0 with the one leaf.
__global__ void k_lb_single(const float4 *leafbox, float4 *nodes4) {
  if (threadIdx.x != 0) return;
  const float4 l4 = leafbox[0], h4 = leafbox[1];
  f3 l = mk3(l4.x, l4.y, l4.z), h = mk3(h4.x, h4.y, h4.z);
  pad_box(l, h);
  nodes4[0] = make_float4(l.x, 1.f, 1.f, 1.f);
  nodes4[1] = make_float4(h.x, -1.f, -1.f, -1.f);
  nodes4[2] = make_float4(l.y, 1.f, 1.f, 1.f);
  nodes4[3] = make_float4(h.y, -1.f, -1.f, -1.f);
  nodes4[4] = make_float4(l.z, 1.f, 1.f, 1.f);
  nodes4[5] = make_float4(h.z, -1.f, -1.f, -1.f);
  nodes4[6] = make_float4(__int_as_float(leaf_code(0u, 1u)), __int_as_float(kEmptyChild),
                          __int_as_float(kEmptyChild), __int_as_float(kEmptyChild));
  nodes4[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

static uint32_t lb_blocks(size_t n) { return (uint32_t)((n + 255) / 256); }

__global__ void k_lb_quantize(const float4 *in, uint32_t n, float4 *out) {
  const uint32_t i = l_tid();
  if (i >= n) return;
  quantize_node(in + 8 * (size_t)i, out + PT_QNODE_F4 * (size_t)i);
}

hipError_t lbvh_quantize(hipStream_t st, const float4 *nodes8, uint32_t n, float4 *nodes_q) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_lb_quantize, dim3(lb_blocks(n)), dim3(256), 0, st, nodes8, n, nodes_q);
  return hipGetLastError();
}

static size_t sort_bytes(uint32_t n) {
  size_t tb = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                           (uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 63);
  return tb;
}

// Scratch layout: every array starts on a 256-byte boundary.
template <class F>
static size_t lbvh_layout(uint32_t n, F place) {
  const size_t m = n > 1 ? n - 1 : 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  place(0, take(2 * (size_t)n * sizeof(float4)));
  place(1, take(2 * (size_t)n * sizeof(float4)));
  place(2, take(2 * m * sizeof(float4)));
  place(3, take((size_t)n * sizeof(uint64_t)));
  place(4, take((size_t)n * sizeof(uint64_t)));
  place(5, take((size_t)n * sizeof(uint32_t)));
  place(6, take((size_t)n * sizeof(uint32_t)));
  place(7, take(m * sizeof(int2)));
  place(8, take(m * sizeof(int2)));
  place(9, take(m * sizeof(int)));
  place(10, take((size_t)n * sizeof(int)));
  place(11, take(m * sizeof(uint32_t)));
  place(12, take(m * sizeof(int2)));
  place(13, take(m * sizeof(int2)));
  place(14, take(12 * sizeof(uint32_t)));
  place(15, take(4 * sizeof(uint32_t)));
  place(16, take(sort_bytes(n)));
  return off;
}

size_t lbvh_scratch_bytes(uint32_t n) {
  return lbvh_layout(n, [](int, size_t) {});
}

void lbvh_carve(void *base, uint32_t n, LbvhScratch &sc) {
  char *b = static_cast<char *>(base);
  lbvh_layout(n, [&](int k, size_t o) {
    void *p = b + o;
    switch (k) {
      case 0: sc.box = (float4 *)p; break;
      case 1: sc.leafbox = (float4 *)p; break;
      case 2: sc.nodebox = (float4 *)p; break;
      case 3: sc.key = (uint64_t *)p; break;
      case 4: sc.key2 = (uint64_t *)p; break;
      case 5: sc.idx = (uint32_t *)p; break;
      case 6: sc.idx2 = (uint32_t *)p; break;
      case 7: sc.child = (int2 *)p; break;
      case 8: sc.range = (int2 *)p; break;
      case 9: sc.parent_int = (int *)p; break;
      case 10: sc.parent_leaf = (int *)p; break;
      case 11: sc.flags = (uint32_t *)p; break;
      case 12: sc.q0 = (int2 *)p; break;
      case 13: sc.q1 = (int2 *)p; break;
      case 14: sc.cb = (uint32_t *)p; break;
      case 15: sc.counts = (uint32_t *)p; break;
      default: sc.sort_tmp = p; break;
    }
  });
  sc.sort_bytes = sort_bytes(n);
}

static float ordered_to_float(uint32_t u) {
  return __builtin_bit_cast(float, (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

hipError_t lbvh_build(hipStream_t st, const float4 *prims, uint32_t n, uint32_t leaf_size, LbvhScratch &sc,
                      float4 *prims_out, float4 *nodes4, LbvhResult &res) {
  hipError_t e;
  hipLaunchKernelGGL(k_lb_init_bounds, dim3(1), dim3(64), 0, st, sc.cb);
  hipLaunchKernelGGL(k_lb_bounds, dim3(lb_blocks(n)), dim3(256), 0, st, prims, n, sc.box, sc.cb);
  hipLaunchKernelGGL(k_lb_morton, dim3(lb_blocks(n)), dim3(256), 0, st, sc.box, n, sc.cb, sc.key, sc.idx);
  size_t tb = sc.sort_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(sc.sort_tmp, tb, sc.key, sc.key2, sc.idx, sc.idx2, (int)n, 0, 63, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_lb_reorder, dim3(lb_blocks(n)), dim3(256), 0, st, prims, sc.box, sc.idx2, n, prims_out,
                     sc.leafbox);
  res.root = 0;
  if (n < 2) {
    hipLaunchKernelGGL(k_lb_single, dim3(1), dim3(64), 0, st, sc.leafbox, nodes4);
    res.nodes4 = 1;
    res.depth4 = 1;
    res.level_end[0] = 1;
  } else {
    e = hipMemsetAsync(sc.flags, 0, (size_t)(n - 1) * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_lb_karras, dim3(lb_blocks(n - 1)), dim3(256), 0, st, sc.key2, (int)n, sc.child,
                       sc.range, sc.parent_int, sc.parent_leaf);
    e = hipMemsetAsync(sc.counts, 0, 4 * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    // the frontier list borrows the first collapse work list (both n - 1 long)
    int *frontier = reinterpret_cast<int *>(sc.q0);
    hipLaunchKernelGGL(k_lb_refit_local, dim3(lb_blocks(n)), dim3(256), 0, st, (int)n, sc.child, sc.range,
                       sc.parent_int, sc.parent_leaf, sc.leafbox, sc.nodebox, sc.flags, leaf_size, frontier,
                       sc.counts);
    uint32_t nf = 0;
    e = hipMemcpyAsync(&nf, sc.counts, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (nf > 0)
      hipLaunchKernelGGL(k_lb_refit_global, dim3(lb_blocks(nf)), dim3(256), 0, st, (int)n, frontier, nf, sc.child,
                         sc.range, sc.parent_int, sc.parent_leaf, sc.leafbox, sc.nodebox, sc.flags, leaf_size);
    hipLaunchKernelGGL(k_lb_root, dim3(1), dim3(64), 0, st, sc.q0, sc.counts, sc.counts + 2);
    // one launch per level of 4-wide nodes; the host reads each level's size
    uint32_t nq = 1, depth = 0, numbered = 1;
    int2 *qa = sc.q0, *qb = sc.q1;
    while (nq > 0) {
      if (depth < 4) res.level_end[depth] = numbered;  // this level's nodes are numbered
      e = hipMemsetAsync(sc.counts + 1, 0, sizeof(uint32_t), st);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(k_lb_collapse, dim3(lb_blocks(nq)), dim3(256), 0, st, qa, nq, sc.child, sc.range,
                         sc.leafbox, sc.nodebox, nodes4, sc.counts + 2, qb, sc.counts + 1);
      e = hipMemcpyAsync(&nq, sc.counts + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
      if (e != hipSuccess) return e;
      e = hipStreamSynchronize(st);
      if (e != hipSuccess) return e;
      numbered += nq;
      ++depth;
      int2 *t = qa;
      qa = qb;
      qb = t;
    }
    res.depth4 = depth;
  }
  uint32_t tail[16];
  e = hipMemcpyAsync(tail, sc.cb, 12 * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  if (n >= 2) {
    e = hipMemcpyAsync(tail + 12, sc.counts, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
  }
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  if (n >= 2) res.nodes4 = tail[14];
  for (int a = 0; a < 3; ++a) {
    res.lo[a] = ordered_to_float(tail[6 + a]);
    res.hi[a] = ordered_to_float(tail[9 + a]);
  }
  return hipGetLastError();
}

}  // namespace pt
