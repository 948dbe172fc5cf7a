// pt_device.h -- traversal, primitive tests, lights, environment and camera
// on the device.
#pragma once
#include "pt_scene.h"
#include "pt_bsdf.h"

namespace pt {

// ---------------------------------------------------------------------
// Sobol direction numbers (Joe & Kuo 2008: d=1 van der Corput; d=2..4 with
// (s, a, m) = (1,0,{1}), (2,1,{1,3}), (3,1,{1,3,1})), built at compile time.
// ---------------------------------------------------------------------
constexpr SobolTable make_sobol() {
  SobolTable t{};
  const int S[4] = {0, 1, 2, 3};
  const int A[4] = {0, 0, 1, 1};
  const uint32_t M[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 3, 0}, {1, 3, 1}};
  for (int k = 0; k < 32; ++k) t.v[0][k] = 1u << (31 - k);
  for (int d = 1; d < 4; ++d) {
    int s = S[d];
    for (int k = 0; k < 32; ++k) {
      if (k < s) {
        t.v[d][k] = M[d][k] << (31 - k);
      } else {
        uint32_t v = t.v[d][k - s] ^ (t.v[d][k - s] >> s);
        for (int j = 1; j < s; ++j)
          if ((A[d] >> (s - 1 - j)) & 1) v ^= t.v[d][k - j];
        t.v[d][k] = v;
      }
    }
  }
  return t;
}
__constant__ static const SobolNibbles kSobol = sobol_nibbles(make_sobol());

// One sampler dimension of path sample k of `pixel` (Halton index g).
PT_D float sample_dim(const DevScene &s, uint32_t sampler, uint32_t seed, uint32_t spp_x, uint32_t spp_y,
                      uint32_t pixel, uint32_t k, uint32_t g, uint32_t dim) {
  switch (sampler) {
    case PT_SAMPLER_SOBOL:
      return sobol_owen(kSobol, k, dim, pt_pixel_seed(pixel, seed));
    case PT_SAMPLER_STRATIFIED:
      return stratified(pt_pixel_seed(pixel, seed), k, spp_x, spp_y, dim);
    case PT_SAMPLER_HALTON_NONE:
      return halton_plain(dim, g);
    case PT_SAMPLER_HALTON_PERM:
      if (dim >= 25u) return halton_plain(dim, g);  // sampler.hpp:61-63
      return halton_perm((int)dim, g, s.perm, s.perm_off);
    default: {
      uint32_t hash = (uint32_t)mix_bits((uint64_t)(1u + (dim << 4)));
      return halton_owen((int)(dim % 25u), g, hash);
    }
  }
}

// ---------------------------------------------------------------------
// Primitive tests.  Triangles: MeshBVHNode::hitTriangle (meshBVH.cpp:85-113)
// on the ray direction as given; spheres and rects normalise the direction
// (primitive.hpp:132, primitive.cpp:38).  All bounds inclusive.  Compiled
// without FMA contraction so that every hit decision is the one the
// reference's unfused FP32 arithmetic makes.
// ---------------------------------------------------------------------
#pragma clang fp contract(off)
PT_D bool hit_tri(float4 A, float4 B, float4 C, f3 o, f3 d, float tmin, float tmax, float &t) {
  f3 p0 = mk3(A.x, A.y, A.z), e1 = mk3(B.x, B.y, B.z), e2 = mk3(C.x, C.y, C.z);
  f3 P = cross3(d, e2);
  float det = dot3(e1, P);
  if (fabsf(det) < 1e-8f) return false;
  float inv = rcp_rn(det);
  f3 T = sub3(o, p0);
  float u = dot3(T, P) * inv;
  if (u < 0.f || u > 1.f) return false;
  f3 Q = cross3(T, e1);
  float v = dot3(d, Q) * inv;
  if (v < 0.f || u + v > 1.f) return false;
  float tt = dot3(e2, Q) * inv;
  if (tt < tmin || tt > tmax) return false;
  t = tt;
  return true;
}
PT_D bool hit_sphere(float4 A, float4 B, f3 o, f3 d, float tmin, float tmax, float &t) {
  f3 dn = norm3_rn(d);
  f3 oc = sub3(o, mk3(A.x, A.y, A.z));
  float r = B.x;
  float a = dot3(dn, dn);
  float b = 2.f * dot3(dn, oc);
  float c = dot3(oc, oc) - r * r;
  float disc = b * b - 4.f * a * c;
  if (disc < 0.f) return false;
  float delta = sqrt_rn(disc);
  float tt = div_rn(-b - delta, 2.f * a);
  if (tt < tmin || tt > tmax) {
    tt = div_rn(-b + delta, 2.f * a);
    if (tt < tmin || tt > tmax) return false;
  }
  t = tt;
  return true;
}
PT_D bool hit_rect(float4 A, float4 B, f3 o, f3 d, float tmin, float tmax, float &t) {
  f3 dn = norm3_rn(d);
  int ax = (int)B.z;
  int b0 = (ax == 0) ? 1 : 0, b1 = (ax == 2) ? 1 : 2;
  float tt = div_rn(A.x - comp3(o, ax), comp3(dn, ax));
  if (tt < tmin || tt > tmax) return false;
  float d0 = comp3(o, b0) + tt * comp3(dn, b0);
  float d1 = comp3(o, b1) + tt * comp3(dn, b1);
  if (d0 < A.y || d0 > A.z || d1 < B.x || d1 > B.y) return false;
  t = tt;
  return true;
}

// One cylinder segment of a Curve (primitives/curve.hpp:47-118): the ray
// against the infinite cylinder about the segment's axis (solveQuadratic,
// utility.hpp:112-124), a root kept when it lies within [tmin, tmax] and its
// projection on the axis within the segment.  Raw direction, as the reference.
PT_D bool hit_curve(float4 A, float4 B, float4 C, f3 o, f3 d, float tmin, float tmax, float &t) {
  const f3 pa = mk3(A.x, A.y, A.z), ax = mk3(B.x, B.y, B.z);
  const float r = B.w, len = C.x;
  const f3 dp = sub3(d, scl3(ax, dot3(d, ax)));
  const f3 de = sub3(o, pa);
  const f3 op = sub3(de, scl3(ax, dot3(de, ax)));
  const float qa = dot3(dp, dp), qb = 2.f * dot3(dp, op), qc = dot3(op, op) - r * r;
  const float disc = qb * qb - 4.f * qa * qc;
  if (disc < 0.f) return false;
  const float sq = sqrt_rn(disc);
  const float q = qb < 0.f ? -0.5f * (qb - sq) : -0.5f * (qb + sq);
  float r0 = div_rn(q, qa), r1 = div_rn(qc, q);
  if (r0 > r1) {
    const float x = r0;
    r0 = r1;
    r1 = x;
  }
  const int nr = (r0 == r1) ? 1 : 2;
  float best = tmax;
  bool found = false;
  for (int k = 0; k < nr; ++k) {
    const float tt = k == 0 ? r0 : r1;
    if (tt < tmin || tt > best) continue;
    const f3 hp = add3(o, scl3(d, tt));
    const float proj = dot3(sub3(hp, pa), ax);
    if (proj < 0.f || proj > len) continue;
    best = tt;
    found = true;
  }
  if (found) t = best;
  return found;
}

PT_D bool hit_prim(const float4 *prims, uint32_t p, f3 o, f3 d, float tmin, float tmax, float &t, uint32_t &id) {
  float4 A = prims[3 * p];
  float4 B = prims[3 * p + 1];
  uint32_t tag = __float_as_uint(A.w);
  id = tag & 0x3fffffffu;
  uint32_t type = tag >> 30;
  if (type == PRIM_TRI) {
    float4 C = prims[3 * p + 2];
    return hit_tri(A, B, C, o, d, tmin, tmax, t);
  }
  if (type == PRIM_SPHERE) return hit_sphere(A, B, o, d, tmin, tmax, t);
  if (type == PRIM_CURVE) {
    float4 C = prims[3 * p + 2];
    return hit_curve(A, B, C, o, d, tmin, tmax, t);
  }
  return hit_rect(A, B, o, d, tmin, tmax, t);
}

// ---------------------------------------------------------------------
// Traversal stack.  Entries [0, lo) live in a private array; with K > 0 the
// newest entries [lo, sp) (at most K) live in an LDS ring at slot i % K, laid
// out slot-major ([K][PT_BLOCK]) so that the lanes of a wave always touch
// distinct banks whatever their depths.  K = 0 keeps everything private.
// The counters are plain locals: kept in a struct beside the arrays they
// were placed in scratch memory too.
// ---------------------------------------------------------------------
#define PT_BLOCK 256
// LDS words per block for a stack of K entries (codes, plus distances); at
// least one word so that the declaration stays valid for K = 0.
// Entries per wave below which a traversal launch gives every wave a fixed
// share instead of sharing a work queue (the Cornell frame's 32k-entry
// partitions: 0.61 ms of closest hit with fixed shares, 0.66 ms queued).
#ifndef PT_QUEUE_MIN
#define PT_QUEUE_MIN 256
#endif
#define PT_TRAV_LDS(K, WITH_T) ((K) > 0 ? (K) * PT_BLOCK * ((WITH_T) ? 2 : 1) : 1)
// Copies the scene's top nodes into the block's LDS (every thread calls it):
// the quantised nodes for any-hit traversal (QN), else the 128-byte nodes.
template <bool QN>
PT_D void stage_top_nodes(const DevScene &s, float4 *top, bool batch) {
  const uint32_t n = batch ? 0u : (QN ? (uint32_t)PT_QNODE_F4 : 8u) * s.top_nodes;
  const float4 *src = QN ? s.qnodes : s.nodes;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) top[i] = src[i];
  __syncthreads();
}
// float4 per node in the array traversal reads (QN: any hit)
template <bool QN>
constexpr int node_f4() { return QN ? PT_QNODE_F4 : 8; }
// Pointers into a block's LDS, typed so that loads through them compile to
// ds_read (a generic pointer that may address LDS or global memory compiles
// to flat loads, which wait on both counters).
typedef __attribute__((address_space(3))) int lds_int;
typedef __attribute__((address_space(3))) float lds_float;
// The four child boxes (SoA) and codes of node nd: a quantised node (QN,
// decode_node) or a 128-byte one.  P: const float4 * (global) or lds_v4 * (LDS).
template <bool QN, class P>
PT_D void load_node(P nd, float (&lx)[4], float (&hx)[4], float (&ly)[4], float (&hy)[4],
                    float (&lz)[4], float (&hz)[4], int4 &ch) {
  if constexpr (QN) {
    decode_node(nd, lx, hx, ly, hy, lz, hz, ch);
  } else {
    const float4 a = ld_f4(nd, 0), b = ld_f4(nd, 1), c = ld_f4(nd, 2), d = ld_f4(nd, 3), e = ld_f4(nd, 4),
                 f = ld_f4(nd, 5), g = ld_f4(nd, 6);
    ch = make_int4((int)pt_float_bits(g.x), (int)pt_float_bits(g.y), (int)pt_float_bits(g.z), (int)pt_float_bits(g.w));
    lx[0] = a.x; lx[1] = a.y; lx[2] = a.z; lx[3] = a.w;
    hx[0] = b.x; hx[1] = b.y; hx[2] = b.z; hx[3] = b.w;
    ly[0] = c.x; ly[1] = c.y; ly[2] = c.z; ly[3] = c.w;
    hy[0] = d.x; hy[1] = d.y; hy[2] = d.z; hy[3] = d.w;
    lz[0] = e.x; lz[1] = e.y; lz[2] = e.z; lz[3] = e.w;
    hz[0] = f.x; hz[1] = f.y; hz[2] = f.z; hz[3] = f.w;
  }
}

// ---------------------------------------------------------------------
// BVH4 traversal.  One 128-byte node per step: the four child boxes are
// tested together, the nearest hit child is visited next and the others are
// pushed far-to-near with their entry distance, so a pop whose box starts
// beyond the current closest hit is skipped.  ANY: stop at the first
// primitive hit in [tmin, tmax] (shadow rays); no ordering needed.  QN: read
// the 64-byte quantised nodes (conservative boxes, same hits) instead of the
// 128-byte ones.
// `lds` is the calling block's PT_TRAV_LDS(K, !ANY) words of LDS.
// ---------------------------------------------------------------------
// Compare-exchange of (distance, code) pairs as selects, so that the network
// is straight-line code (v_cndmask) rather than exec-masked branches.
#define PT_CSWAP(ta, ca, tb_, cb)                                   \
  {                                                                 \
    const bool _s = tb_ < ta;                                       \
    const float _lo = _s ? tb_ : ta, _hi = _s ? ta : tb_;           \
    const int _cl = _s ? cb : ca, _ch = _s ? ca : cb;               \
    ta = _lo; tb_ = _hi; ca = _cl; cb = _ch;                        \
  }
template <bool ANY, bool QN, bool COUNT, int K>
PT_D bool traverse(const DevScene &s, int *lds, f3 o, f3 d, float tmin, float tmax, float &t_out, int &prim_out,
                   uint32_t &n_nodes, uint32_t &n_prims) {
  const float ix = rcp_fast(d.x), iy = rcp_fast(d.y), iz = rcp_fast(d.z);
  const float oix = o.x * ix, oiy = o.y * iy, oiz = o.z * iz;
  float tb = tmax;
  int best = -1;
  constexpr bool WT = !ANY;
  int pc[PT_STACK4];
  float ptv[WT ? PT_STACK4 : 1];
  int *const lc = lds + threadIdx.x;
  float *const lt = reinterpret_cast<float *>(lds + K * PT_BLOCK) + threadIdx.x;
  int sp = 0, lo = 0;
  auto push = [&](int c, float t) {
    if constexpr (K > 0) {
      if (sp - lo == K) {  // ring full: the oldest entry goes to the private array
        const int s0 = (lo & (K - 1)) * PT_BLOCK;
        pc[lo] = lc[s0];
        if (WT) ptv[lo] = lt[s0];
        ++lo;
      }
      const int s1 = (sp & (K - 1)) * PT_BLOCK;
      lc[s1] = c;
      if (WT) lt[s1] = t;
    } else {
      pc[sp] = c;
      if (WT) ptv[sp] = t;
    }
    ++sp;
  };
  auto pop = [&](int &c, float &t) {
    --sp;
    if constexpr (K > 0) {
      if (sp >= lo) {
        const int s1 = (sp & (K - 1)) * PT_BLOCK;
        c = lc[s1];
        t = WT ? lt[s1] : 0.f;
        return;
      }
      lo = sp;
    }
    c = pc[sp];
    t = WT ? ptv[sp] : 0.f;
  };
  int cur = s.root;
  while (true) {
    if (cur >= 0) {
      float lxs[4], hxs[4], lys[4], hys[4], lzs[4], hzs[4];
      int4 ch;
      load_node<QN>((QN ? s.qnodes : s.nodes) + node_f4<QN>() * (size_t)cur, lxs, hxs, lys, hys, lzs, hzs, ch);
      if (COUNT) ++n_nodes;
      float tn[4], tf[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float a0 = fmaf(lxs[k], ix, -oix), a1 = fmaf(hxs[k], ix, -oix);
        float b0 = fmaf(lys[k], iy, -oiy), b1 = fmaf(hys[k], iy, -oiy);
        float c0 = fmaf(lzs[k], iz, -oiz), c1 = fmaf(hzs[k], iz, -oiz);
        tn[k] = fmaxf(fmaxf(fminf(a0, a1), fminf(b0, b1)), fmaxf(fminf(c0, c1), tmin));
        tf[k] = fminf(fminf(fmaxf(a0, a1), fmaxf(b0, b1)), fminf(fmaxf(c0, c1), tb));
      }
      const int cs[4] = {ch.x, ch.y, ch.z, ch.w};
      float t0, t1, t2, t3;
      int c0 = cs[0], c1 = cs[1], c2 = cs[2], c3 = cs[3];
      int nh = 0;
      {
        bool h0 = tn[0] <= tf[0] && cs[0] != kEmptyChild;
        bool h1 = tn[1] <= tf[1] && cs[1] != kEmptyChild;
        bool h2 = tn[2] <= tf[2] && cs[2] != kEmptyChild;
        bool h3 = tn[3] <= tf[3] && cs[3] != kEmptyChild;
        t0 = h0 ? tn[0] : INFINITY;
        t1 = h1 ? tn[1] : INFINITY;
        t2 = h2 ? tn[2] : INFINITY;
        t3 = h3 ? tn[3] : INFINITY;
        nh = (int)h0 + (int)h1 + (int)h2 + (int)h3;
      }
      if (nh > 0) {
        // sorting network on the entry distance; misses (inf) sink to the end
        PT_CSWAP(t0, c0, t1, c1);
        PT_CSWAP(t2, c2, t3, c3);
        PT_CSWAP(t0, c0, t2, c2);
        PT_CSWAP(t1, c1, t3, c3);
        PT_CSWAP(t1, c1, t2, c2);
        if (nh >= 4) push(c3, t3);
        if (nh >= 3) push(c2, t2);
        if (nh >= 2) push(c1, t1);
        cur = c0;
        continue;
      }
    } else {
      uint32_t st = leaf_start(cur), cn = leaf_count(cur);
      for (uint32_t p = st; p < st + cn; ++p) {
        float t;
        uint32_t id;
        if (COUNT) ++n_prims;
        if (hit_prim(s.prims, p, o, d, tmin, tb, t, id)) {
          tb = t;
          best = (int)id;
          if (ANY) {
            t_out = t;
            prim_out = best;
            return true;
          }
        }
      }
    }
    // pop the next subtree whose box starts no farther than the closest hit
    bool found = false;
    while (sp > 0) {
      int c;
      float te;
      pop(c, te);
      if (ANY || te <= tb) {
        cur = c;
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  t_out = tb;
  prim_out = best;
  return best >= 0;
}

// ---------------------------------------------------------------------
// Persistent traversal with per-lane ray refill and speculative while-while
// loops (Aila & Laine, HPG 2009, on 64-wide waves).  The rays [0, n) of a
// queue partition are shared by its nwaves waves: wave wid first takes a
// fixed contiguous share of the first ST/8 of them, then chunks of CH
// consecutive entries of the rest through the partition's counter work[0]
// (on the 1.3M-triangle glass frame, ST = 0 against fixed shares alone:
// closest hit 30.6 -> 25.9 ms, any hit 17.5 -> 15.8 ms; ST = 4: 26.0, 17.0).
// Each iteration runs a node phase and a leaf phase (below), and lanes whose
// rays are finished take the next rays of the wave's range (a new chunk when
// it runs dry) once REFILL of them are idle, so lanes do not idle while the
// wave's longest ray finishes, and waves that drew cheap rays take more of
// the rest instead of idling while others finish their share; the fixed
// shares keep the counter's round trips to the tail of the launch.  The
// per-ray results equal traverse<>'s (the closest hit, or whether any hit
// exists) up to the winner of exact ties: a postponed leaf is tested after
// some of the nodes traverse<> would visit after it, which changes the order
// of tests and culls fewer subtrees.  work[0] must be zero at launch: a kernel
// between two launches rewinds it (a count of finished waves that lets the
// last one rewind it needs a device-scope release per wave, which writes back
// the L2 and cost a launch of the Cornell frame 0.13 ms).  fetch(i, o, d, tmin,
// tmax, scene) loads ray i (and, batched, points scene at the ray's instance);
// finish(i, hit, t, prim) stores it.
// ---------------------------------------------------------------------
template <bool ANY, bool QN, bool COUNT, int K, int REFILL, int ST, int CH, bool BATCH, class Fetch, class Finish>
PT_D void trace_queue(const DevScene &s, int *lds, const float4 *top, uint32_t n, uint32_t *work, uint32_t wid,
                      uint32_t nwaves, Fetch fetch, Finish finish, uint32_t &n_nodes, uint32_t &n_prims) {
  // nodes [0, ntop) are read from the block's LDS copy `top` (single scenes)
  const uint32_t ntop = BATCH ? 0u : s.top_nodes;
  // BATCH: fetch names each ray's scene, whose tree the lane then walks
  // QN: walk the quantised nodes (any hit over a large tree)
  const float4 *const snodes = QN ? s.qnodes : s.nodes;
  constexpr int NF4 = node_f4<QN>();
  const float4 *lane_nodes = snodes, *lane_prims = s.prims;
  int lane_root = s.root;
  constexpr bool WT = !ANY;
  int pc[PT_STACK4];
  float ptv[WT ? PT_STACK4 : 1];
  lds_int *const lc = (lds_int *)lds + threadIdx.x;
  lds_float *const lt = (lds_float *)(lds + K * PT_BLOCK) + threadIdx.x;
  const uint32_t lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  // the fixed shares, then the entries [dyn, n) that the counter hands out;
  // below PT_QUEUE_MIN entries per wave every entry is in a fixed share (a
  // launch that short would wait on the counter's round trip)
  const bool shared = n >= (uint32_t)PT_QUEUE_MIN * nwaves;
  const uint32_t per = shared ? (uint32_t)((uint64_t)n * (uint32_t)ST / (8ull * nwaves)) : (n + nwaves - 1u) / nwaves;
  const uint32_t dyn = shared ? per * nwaves : n;
  uint32_t next = min(n, wid * per), wend = min(n, next + per);  // the wave's unclaimed entries
  bool more = dyn < n;  // wave-uniform: the counter may still hand out entries
  int ray = -1;
  f3 o = s3(0.f), d = s3(0.f);
  float ix = 0.f, iy = 0.f, iz = 0.f, oix = 0.f, oiy = 0.f, oiz = 0.f, tmin = 0.f, tb = 0.f;
  int best = -1, cur = 0, sp = 0, lo = 0;
  auto push = [&](int c, float t) {
    if constexpr (K > 0) {
      if (sp - lo == K) {
        const int s0 = (lo & (K - 1)) * PT_BLOCK;
        pc[lo] = lc[s0];
        if (WT) ptv[lo] = lt[s0];
        ++lo;
      }
      const int s1 = (sp & (K - 1)) * PT_BLOCK;
      lc[s1] = c;
      if (WT) lt[s1] = t;
    } else {
      pc[sp] = c;
      if (WT) ptv[sp] = t;
    }
    ++sp;
  };
  auto pop = [&](int &c, float &t) {
    --sp;
    if constexpr (K > 0) {
      if (sp >= lo) {
        const int s1 = (sp & (K - 1)) * PT_BLOCK;
        c = lc[s1];
        t = WT ? lt[s1] : 0.f;
        return;
      }
      lo = sp;
    }
    c = pc[sp];
    t = WT ? ptv[sp] : 0.f;
  };
  // nearest stack entry whose box starts no farther than the closest hit
  auto pop_next = [&](int &c) -> bool {
    while (sp > 0) {
      int cc;
      float te;
      pop(cc, te);
      if (ANY || te <= tb) {
        c = cc;
        return true;
      }
    }
    return false;
  };
  bool have_cur = false, have_leaf = false;
  int leaf = 0;
  uint32_t lq = 0, lend = 0;  // any hit: the primitive cursor of the current leaf
  // one inner-node step: test the 4 children, go to the nearest hit one and
  // push the others far to near, else pop
  auto node_step = [&]() {
    float lxs[4], hxs[4], lys[4], hys[4], lzs[4], hzs[4];
    int4 ch;
    if constexpr (QN) {
      // one decode on a generic pointer: a decode per address space costs the
      // any-hit kernel spills (shadow rays 23.3 -> 35.5 ms on the 1.3M-triangle frame)
      load_node<QN>(((uint32_t)cur < ntop ? top : (BATCH ? lane_nodes : snodes)) + NF4 * (size_t)cur, lxs, hxs, lys,
                    hys, lzs, hzs, ch);
    } else if ((uint32_t)cur < ntop) {
      load_node<QN>((lds_v4 *)top + NF4 * cur, lxs, hxs, lys, hys, lzs, hzs, ch);
    } else {
      load_node<QN>((BATCH ? lane_nodes : snodes) + NF4 * (size_t)cur, lxs, hxs, lys, hys, lzs, hzs, ch);
    }
    if (COUNT) ++n_nodes;
    float tn[4], tf[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float a0 = fmaf(lxs[k], ix, -oix), a1 = fmaf(hxs[k], ix, -oix);
      float b0 = fmaf(lys[k], iy, -oiy), b1 = fmaf(hys[k], iy, -oiy);
      float c0 = fmaf(lzs[k], iz, -oiz), c1 = fmaf(hzs[k], iz, -oiz);
      tn[k] = fmaxf(fmaxf(fminf(a0, a1), fminf(b0, b1)), fmaxf(fminf(c0, c1), tmin));
      tf[k] = fminf(fminf(fmaxf(a0, a1), fmaxf(b0, b1)), fminf(fmaxf(c0, c1), tb));
    }
    const int cs[4] = {ch.x, ch.y, ch.z, ch.w};
    int c0 = cs[0], c1 = cs[1], c2 = cs[2], c3 = cs[3];
    const bool h0 = tn[0] <= tf[0] && cs[0] != kEmptyChild;
    const bool h1 = tn[1] <= tf[1] && cs[1] != kEmptyChild;
    const bool h2 = tn[2] <= tf[2] && cs[2] != kEmptyChild;
    const bool h3 = tn[3] <= tf[3] && cs[3] != kEmptyChild;
    float t0 = h0 ? tn[0] : INFINITY, t1 = h1 ? tn[1] : INFINITY;
    float t2 = h2 ? tn[2] : INFINITY, t3 = h3 ? tn[3] : INFINITY;
    const int nh = (int)h0 + (int)h1 + (int)h2 + (int)h3;
    if (ANY && nh > 0) {
      // any hit: every order finds an occluder when there is one; visit the
      // first hit child in slot order and push the others unsorted
      if (h3 && nh >= 2) push(c3, 0.f);
      if (h2 && (h0 || h1)) push(c2, 0.f);
      if (h1 && h0) push(c1, 0.f);
      cur = h0 ? c0 : (h1 ? c1 : (h2 ? c2 : c3));
    } else if (nh > 0) {
      PT_CSWAP(t0, c0, t1, c1);
      PT_CSWAP(t2, c2, t3, c3);
      PT_CSWAP(t0, c0, t2, c2);
      PT_CSWAP(t1, c1, t3, c3);
      PT_CSWAP(t1, c1, t2, c2);
      if (nh >= 4) push(c3, t3);
      if (nh >= 3) push(c2, t2);
      if (nh >= 2) push(c1, t1);
      cur = c0;
    } else {
      have_cur = pop_next(cur);
    }
  };
  while (true) {
    // refill the lanes that have no ray once at least REFILL of them are idle
    // (or all are): waiting for a group keeps the new rays of a wave adjacent
    // (wave-uniform bookkeeping); the rest of the chunk first, then a new one
    const unsigned long long idle = __ballot(ray < 0);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if ((next < wend || more) && (n_idle >= (uint32_t)REFILL || n_idle == 64u)) {
      const uint32_t rem = wend - next;
      uint32_t g = 0u, gn = 0u;
      if (rem < n_idle && more) {
        uint32_t v = 0u;
        if (lane == (uint32_t)__builtin_amdgcn_readfirstlane((int)lane)) v = atomicAdd(work, (uint32_t)CH);
        g = dyn + (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
        gn = g < n ? min((uint32_t)CH, n - g) : 0u;
        more = g + (uint32_t)CH < n;
      }
      if (ray < 0) {
        const uint32_t k = (uint32_t)__popcll(idle & below);
        const bool got = k < rem || k - rem < gn;
        if (got) {
          const uint32_t mine = k < rem ? next + k : g + (k - rem);
          ray = (int)mine;
          float tmax;
          const DevScene *sc = &s;
          fetch(mine, o, d, tmin, tmax, sc);
          if (BATCH) {
            lane_nodes = QN ? sc->qnodes : sc->nodes;
            lane_prims = sc->prims;
            lane_root = sc->root;
          }
          ix = rcp_fast(d.x); iy = rcp_fast(d.y); iz = rcp_fast(d.z);
          oix = o.x * ix; oiy = o.y * iy; oiz = o.z * iz;
          tb = tmax;
          best = -1;
          cur = BATCH ? lane_root : s.root;
          have_cur = true;
          have_leaf = false;
          lq = lend = 0;
          sp = 0;
          lo = 0;
        }
      }
      if (rem >= n_idle) {
        next += n_idle;
      } else {
        next = g + min(gn, n_idle - rem);
        wend = g + gn;
      }
    }
    if (__ballot(ray >= 0) == 0ull) break;
    if (ray < 0) continue;
    if constexpr (ANY) {
      // Any hit: one node or one leaf per iteration.  One hit ends the ray,
      // so a leaf is tested as soon as it is reached (the speculative loops
      // below cost the shadow rays 8% on the 1.3M-triangle frame, plain
      // while-while loops 26%).
      if (cur >= 0) {
        node_step();
      } else {
        // one primitive of the current leaf per iteration, so that a lane in a
        // leaf holds the lanes stepping through nodes back by one test only
        if (lq >= lend) {
          lq = leaf_start(cur);
          lend = lq + leaf_count(cur);
        }
        float t;
        uint32_t id;
        if (COUNT) ++n_prims;
        if (hit_prim(BATCH ? lane_prims : s.prims, lq, o, d, tmin, tb, t, id)) {
          tb = t;
          best = (int)id;
          have_cur = false;
        } else if (++lq >= lend) {
          have_cur = pop_next(cur);
        }
      }
      if (!have_cur) {
        finish((uint32_t)ray, best >= 0, tb, best);
        ray = -1;
      }
      continue;
    }
    // Node phase (speculative while-while, Aila & Laine 2009): lanes step
    // through inner nodes; the first leaf a lane reaches is postponed and
    // the lane goes on with its stack, until every lane holds a leaf or has
    // no inner node left, so that the node steps and the leaf tests each run
    // with most lanes of the wave at once (closest hit: 46.8 -> 44.2 ms on
    // the 1.3M-triangle frame).
    while (true) {
      const bool step = have_cur && cur >= 0;
      if (__ballot(step) == 0ull) break;
      if (step) node_step();
      // postpone the first leaf reached
      if (have_cur && cur < 0 && !have_leaf) {
        leaf = cur;
        have_leaf = true;
        have_cur = pop_next(cur);
      }
      // stop once no lane is still looking for a leaf among inner nodes
      if (__ballot(have_cur && cur >= 0 && !have_leaf) == 0ull) break;
    }
    // Leaf phase: the postponed leaf, then leaves the stack yields next.
    bool done = false;
    while (true) {
      if (__ballot(have_leaf) == 0ull) break;
      if (have_leaf) {
        const uint32_t st = leaf_start(leaf), cn = leaf_count(leaf);
        for (uint32_t q = st; q < st + cn; ++q) {
          float t;
          uint32_t id;
          if (COUNT) ++n_prims;
          if (hit_prim(BATCH ? lane_prims : s.prims, q, o, d, tmin, tb, t, id)) {
            tb = t;
            best = (int)id;
            if (ANY) {
              done = true;
              break;
            }
          }
        }
        if (done) {
          have_leaf = have_cur = false;
        } else if (have_cur && cur < 0) {
          leaf = cur;
          have_cur = pop_next(cur);
        } else {
          have_leaf = false;
        }
      }
    }
    if (!have_cur && !have_leaf) {
      finish((uint32_t)ray, best >= 0, tb, best);
      ray = -1;
    }
  }
}
#undef PT_CSWAP

// ---------------------------------------------------------------------
// Shading geometry.  HitRecord::set_normal (materialBase.hpp:48-52) and
// the wrapper flags (FlipFace toggles is_inside; Translate/Rotate/Scale
// force it true).
// ---------------------------------------------------------------------
struct SurfHit {
  f3 p;
  f3 ng;  // geometric normal before flipping (light-facing for rects)
  ShadePoint sp;
  uint32_t mat;
  int light;
};
// Interpolated shading normal of a triangle hit (meshBVH.cpp:116-130): the
// barycentrics (u, v) of MeshBVHNode::hitTriangle recomputed from the same
// ray with the same arithmetic as hit_tri, then normalize((1-u-v) n0 + u n1 +
// v n2), turned to the geometric normal's side.
#pragma clang fp contract(off)
PT_D f3 smooth_normal(const float4 *r, f3 o, f3 d, f3 ng) {
  const float4 a = r[0], b = r[1], c = r[2], e = r[3], f = r[4];
  f3 p0 = mk3(a.x, a.y, a.z), e1 = mk3(b.x, b.y, b.z), e2 = mk3(c.x, c.y, c.z);
  f3 P = cross3(d, e2);
  float inv = rcp_rn(dot3(e1, P));
  f3 T = sub3(o, p0);
  float u = dot3(T, P) * inv;
  f3 Q = cross3(T, e1);
  float v = dot3(d, Q) * inv;
  f3 n0 = mk3(a.w, b.w, c.w), n1 = mk3(e.x, e.y, e.z), n2 = mk3(e.w, f.x, f.y);
  f3 ns = norm3_rn(add3(add3(scl3(n0, 1.f - u - v), scl3(n1, u)), scl3(n2, v)));
  return dot3(ns, ng) < 0.f ? neg3(ns) : ns;
}

PT_D SurfHit surface(const DevScene &s, int prim, f3 o, f3 d, float t) {
  SurfHit h;
  uint2 info = prim_info(s, prim);
  float4 g = prim_geom(s, prim);
  uint32_t type = (info.x >> 24) & 3u, fl = info.x >> 28;
  h.mat = info.x & 0xffffffu;
  h.light = (type == PRIM_TRI || type == PRIM_CURVE) ? -1 : (int)info.y;
  f3 n;
  f3 tan = s3(0.f);
  if (type == PRIM_CURVE) {
    // Curve::intersect's hit record (curve.hpp:101-129): the tangent from the
    // Bezier derivatives at the segment's ends, the normal from the axis to the
    // hit point, perpendicular to the tangent
    const float4 *cg = s.curve_geom + 4 * (size_t)info.y;
    const f3 pa = mk3(cg[0].x, cg[0].y, cg[0].z), ax = mk3(cg[1].x, cg[1].y, cg[1].z);
    const f3 d0 = mk3(cg[2].x, cg[2].y, cg[2].z), d1 = mk3(cg[3].x, cg[3].y, cg[3].z);
    h.p = add3(o, scl3(d, t));
    const float proj = dot3(sub3(h.p, pa), ax);
    const float uu = div_rn(proj, cg[0].w);
    tan = norm3(add3(d0, scl3(sub3(d1, d0), uu)));
    const f3 v = sub3(h.p, add3(pa, scl3(ax, proj)));
    n = norm3(sub3(v, scl3(tan, dot3(v, tan))));
  } else if (type == PRIM_TRI) {
    h.p = add3(o, scl3(d, t));
    n = mk3(g.x, g.y, g.z);
    h.ng = n;
    if (info.y != 0xffffffffu) n = smooth_normal(s.tri_smooth + 5 * (size_t)info.y, o, d, n);
  } else if (type == PRIM_SPHERE) {
    h.p = add3(o, scl3(norm3(d), t));
    n = norm3(sub3(h.p, mk3(g.x, g.y, g.z)));
  } else {
    h.p = add3(o, scl3(norm3(d), t));
    int ax = (int)g.x;
    n = ax == 0 ? mk3(1.f, 0.f, 0.f) : (ax == 1 ? mk3(0.f, 1.f, 0.f) : mk3(0.f, 0.f, 1.f));
  }
  if (type != PRIM_TRI) h.ng = n;
  bool inside = dot3(d, n) < 0.f;
  h.sp.n = inside ? n : neg3(n);
  if (fl & PT_PRIM_FORCE_INSIDE) inside = true;
  if (fl & PT_PRIM_FLIP_FACE) inside = !inside;
  h.sp.inside = inside;
  h.sp.t = tan;
  return h;
}

// ---------------------------------------------------------------------
// Textures.  ImageTexture::getTexelColor, bilinear (texture.cpp:111-142):
// the lower texel index is (uint)(u w), the fraction taken before it is
// clamped to w-2 (h-2): in the last cell the weights are those of the cell
// one texel back, and u = 1 reads texel w-2; negative coordinates convert
// to index 0 (and extrapolate), as a saturating conversion gives.  CheckerTexture (texture.cpp:165-183): colour 1 where
// sin(u scale) sin(v scale) < 0.
// ---------------------------------------------------------------------
PT_D f3 tex_texel(const float *tx, uint32_t w, uint32_t x, uint32_t y) {
  const float *c = tx + 3 * ((size_t)y * w + x);
  return mk3(c[0], c[1], c[2]);
}
PT_D f3 texture_lookup(const DevScene &s, uint32_t id, float u, float v) {
  const float4 a = s.tex_desc[2 * id];
  if (__float_as_uint(a.x) == TEX_CHECKER) {
    const float4 b = s.tex_desc[2 * id + 1];
    const float sv = sinf(u * a.y) * sinf(v * a.y);
    return sv < 0.f ? mk3(a.z, a.w, b.x) : mk3(b.y, b.z, b.w);
  }
  const uint32_t w = __float_as_uint(a.y), h = __float_as_uint(a.z);
  const float *tx = s.tex + 3 * (size_t)__float_as_uint(a.w);
  const float fx = u * (float)w, fy = v * (float)h;
  uint32_t x = fx > 0.f ? (uint32_t)fx : 0u, y = fy > 0.f ? (uint32_t)fy : 0u;
  const float x1 = fx - (float)x, y1 = fy - (float)y;
  x = min(x, w - 2u);
  y = min(y, h - 2u);
  const float x2 = 1.0f - x1, y2 = 1.0f - y1;
  const f3 c1 = tex_texel(tx, w, x, y), c2 = tex_texel(tx, w, x + 1, y);
  const f3 c3 = tex_texel(tx, w, x, y + 1), c4 = tex_texel(tx, w, x + 1, y + 1);
  const f3 c12 = add3(scl3(c1, x2), scl3(c2, x1));
  const f3 c34 = add3(scl3(c3, x2), scl3(c4, x1));
  return add3(scl3(c12, y2), scl3(c34, y1));
}
// HitRecord::texCoords of a hit on primitive prim at h.p:
//   sphere  (1 - (phi + pi) / 2pi, (theta + pi/2) / pi) of the unit vector to
//           the hit (phi = atan2(z, x), theta = asin(y)), u shifted by 0.75
//           modulo 1 (primitive.hpp:157-163)
//   rect    the hit's position within the extents (primitive.cpp:55-57)
//   triangle (1-u-v) uv0 + u uv1 + v uv2 with the hit's barycentrics
//           (meshBVH.cpp:133-138), or without uvs TriangleMesh's planar
//           fallback: the two axes the normal points least along, scaled
//           to the mesh box and clamped to [0, 1] (triangleMesh.cpp:38-51)
//   curve   (0, 0): Curve sets none
#pragma clang fp contract(off)
PT_D float2 hit_uv(const DevScene &s, int prim, const SurfHit &h, f3 o, f3 d) {
  const uint2 info = prim_info(s, prim);
  const uint32_t type = (info.x >> 24) & 3u;
  const float4 g = prim_geom(s, prim);
  if (type == PRIM_SPHERE) {
    const f3 p = norm3(sub3(h.p, mk3(g.x, g.y, g.z)));
    const float phi = atan2f(p.z, p.x), theta = asinf(p.y);
    const double pi = 3.14159265358979323846;
    float u = (float)(1.0 - ((double)phi + pi) / (2.0 * pi));
    const float v = (float)(((double)theta + pi / 2.0) / pi);
    u = fmodf(u + 0.75f, 1.0f);
    return make_float2(u, v);
  }
  const float4 pt = s.prim_tex[prim];
  if (type == PRIM_RECT) {
    const int ax = (int)g.x;
    const float d0 = ax == 0 ? h.p.y : h.p.x, d1 = ax == 2 ? h.p.y : h.p.z;
    return make_float2((d0 - pt.x) / (pt.y - pt.x), (d1 - pt.z) / (pt.w - pt.z));
  }
  if (type != PRIM_TRI) return make_float2(0.f, 0.f);
  const uint32_t rec = __float_as_uint(pt.y);
  if (rec != 0xffffffffu) {
    const float4 *r = s.tri_uv + 4 * (size_t)rec;
    const float4 a = r[0], b = r[1], c = r[2], e = r[3];
    const f3 p0 = mk3(a.x, a.y, a.z), e1 = mk3(b.x, b.y, b.z), e2 = mk3(c.x, c.y, c.z);
    const f3 P = cross3(d, e2);
    const float inv = rcp_rn(dot3(e1, P));
    const f3 T = sub3(o, p0);
    const float bu = dot3(T, P) * inv;
    const f3 Q = cross3(T, e1);
    const float bv = dot3(d, Q) * inv;
    const float w0 = 1.f - bu - bv;
    return make_float2(w0 * a.w + bu * c.w + bv * e.y, w0 * b.w + bu * e.x + bv * e.z);
  }
  const uint32_t mesh = __float_as_uint(pt.x);
  const float4 lo = s.mesh_box[2 * mesh], hi = s.mesh_box[2 * mesh + 1];
  const f3 n = mk3(fabsf(h.sp.n.x), fabsf(h.sp.n.y), fabsf(h.sp.n.z));
  const int au = (n.x > n.y && n.x > n.z) ? 1 : 0;
  const int av = au == 1 ? 2 : (n.y > n.z ? 2 : 1);
  const float pl[3] = {h.p.x, h.p.y, h.p.z}, ml[3] = {lo.x, lo.y, lo.z}, mh[3] = {hi.x, hi.y, hi.z};
  const float u = (pl[au] - ml[au]) / (mh[au] - ml[au]), v = (pl[av] - ml[av]) / (mh[av] - ml[av]);
  return make_float2(fminf(fmaxf(u, 0.f), 1.f), fminf(fmaxf(v, 0.f), 1.f));
}
// surface() plus, for a textured material (material sets with
// kMatTextured), its colour at the hit's texture coordinates.
template <uint32_t MS>
PT_D SurfHit surface_tex(const DevScene &s, int prim, f3 o, f3 d, float t) {
  SurfHit h = surface(s, prim, o, d, t);
  if constexpr (PT_IN(MS, 31)) {
    const uint32_t fl = s.mats[h.mat].flags;
    if (fl & 1u) {
      const float2 uv = hit_uv(s, prim, h, o, d);
      h.sp.tex = texture_lookup(s, fl >> 8, uv.x, uv.y);
      h.sp.has_tex = true;
    }
  }
  return h;
}

// ---------------------------------------------------------------------
// Environment.  HDRITexture::getTexelColor (texture.hpp:105-115) and the
// bilinear ImageTexture lookup with index clamp to size-2 taken after the
// fractions (texture.cpp:111-142).
// ---------------------------------------------------------------------
// HenyeyGreenstein (medium/phaseMath.hpp:7-15), PBRT's (wo, wi) sign convention.
PT_D float hg_phase(float cos_theta, float g) {
  const float upper = 1.f - g * g;
  float lower = 1.f + g * g + 2.f * g * cos_theta;
  lower *= sqrtf(fmaxf(0.f, lower));
  return upper * 0.0795774715f / lower;  // 1 / (4 pi) rounded to float
}

PT_D f3 env_texel(const DevScene &s, uint32_t x, uint32_t y) {
  const float *c = s.env + 3 * ((size_t)y * s.env_w + x);
  return mk3(c[0], c[1], c[2]);
}
PT_D f3 env_lookup(const DevScene &s, f3 dir) {
  f3 dn = norm3(dir);
  float theta = acosf(-dn.y);
  float phi = atan2f(-dn.z, dn.x) + PT_PI;
  float u = phi / (2.f * PT_PI), v = theta / PT_PI;
  float fx = u * (float)s.env_w, fy = v * (float)s.env_h;
  uint32_t x = (uint32_t)fx, y = (uint32_t)fy;
  float x1 = fx - (float)x, y1 = fy - (float)y;
  x = min(x, (uint32_t)(s.env_w - 2));
  y = min(y, (uint32_t)(s.env_h - 2));
  float x2 = 1.f - x1, y2 = 1.f - y1;
  f3 c1 = env_texel(s, x, y), c2 = env_texel(s, x + 1, y);
  f3 c3 = env_texel(s, x, y + 1), c4 = env_texel(s, x + 1, y + 1);
  f3 c12 = add3(scl3(c1, x2), scl3(c2, x1));
  f3 c34 = add3(scl3(c3, x2), scl3(c4, x1));
  return scl3(add3(scl3(c12, y2), scl3(c34, y1)), s.env_scale);
}
// ImageTexture::getTexelColor at texture coordinates (u, v) (texture.cpp:
// 111-142), times the environment scale: EnvironmentLight::texture->
// getTexelColor(offset) of the reservoir estimator.
PT_D f3 env_lookup_uv(const DevScene &s, float u, float v) {
  float fx = u * (float)s.env_w, fy = v * (float)s.env_h;
  uint32_t x = (uint32_t)fx, y = (uint32_t)fy;
  float x1 = fx - (float)x, y1 = fy - (float)y;
  x = min(x, (uint32_t)(s.env_w - 2));
  y = min(y, (uint32_t)(s.env_h - 2));
  float x2 = 1.f - x1, y2 = 1.f - y1;
  f3 c1 = env_texel(s, x, y), c2 = env_texel(s, x + 1, y);
  f3 c3 = env_texel(s, x, y + 1), c4 = env_texel(s, x + 1, y + 1);
  f3 c12 = add3(scl3(c1, x2), scl3(c2, x1));
  f3 c34 = add3(scl3(c3, x2), scl3(c4, x1));
  return scl3(add3(scl3(c12, y2), scl3(c34, y1)), s.env_scale);
}
// uniformSampleSphere (utils/random.hpp:7-15)
PT_D f3 uniform_sphere(float u1, float u2) {
  float z = 1.f - 2.f * u1;
  float r = sqrtf(fmaxf(0.f, 1.f - z * z));
  float ph = 2.f * PT_PI * u2;
  return mk3(r * cosf(ph), r * sinf(ph), z);
}
PT_D uint32_t find_interval(const float *cdf, uint32_t n, float u) {
  uint32_t lo = 0, hi = n;
  while (lo + 1 < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid; else hi = mid;
  }
  return lo;
}
// find_interval with the search narrowed by a bucket table (the cutpoint
// method of Chen and Asau, 1974): bucket b of G holds an index range that
// contains the answer for every u of the bucket, so the result is the same
// as find_interval's (the largest i < n with cdf[i] <= u) in a few steps
// instead of log2(n) dependent loads.
PT_D uint32_t find_interval_guided(const float *cdf, uint32_t n, float u, const uint2 *guide, uint32_t G) {
  const uint2 r = guide[min(G - 1u, (uint32_t)(u * (float)G))];
  uint32_t lo = r.x, hi = min(n, r.y + 1u);
  while (lo + 1 < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid; else hi = mid;
  }
  return lo;
}
// 2-D piecewise-constant environment sampling (the reference samples the
// environment uniformly, texture.hpp:120-131; this is the MIS estimator's).
// uv (optional) receives the sampled texture coordinates, which the returned
// direction maps back to, so that the radiance along it is read there
// without the inverse mapping; x < 0 where the map is sampled uniformly.
PT_D f3 env_sample(const DevScene &s, float u0, float u1, float &pdf, float2 *uv = nullptr) {
  uint32_t W = s.env_w, H = s.env_h;
  if (!(s.env_integral > 0.f)) {
    pdf = 1.f / (4.f * PT_PI);
    if (uv) *uv = make_float2(-1.f, 0.f);
    return uniform_sphere(u0, u1);
  }
  uint32_t y = find_interval_guided(s.env_mcdf, H, u1, s.env_mguide, H);
  float dy = u1 - s.env_mcdf[y], wy = s.env_mcdf[y + 1] - s.env_mcdf[y];
  if (wy > 0.f) dy /= wy;
  const float *cdf = s.env_cdf + (size_t)y * (W + 1);
  uint32_t x = find_interval_guided(cdf, W, u0, s.env_guide + (size_t)y * W, W);
  float dx = u0 - cdf[x], wx = cdf[x + 1] - cdf[x];
  if (wx > 0.f) dx /= wx;
  float u = ((float)x + dx) / (float)W, v = ((float)y + dy) / (float)H;
  if (uv) *uv = make_float2(u, v);
  float th = v * PT_PI, ph = u * 2.f * PT_PI;
  float st = sinf(th);
  float f = s.env_func[(size_t)y * W + x];
  pdf = (st > 0.f) ? f / s.env_integral / (2.f * PT_PI * PT_PI * st) : 0.f;
  return mk3(-st * cosf(ph), -cosf(th), st * sinf(ph));
}
PT_D float env_pdf(const DevScene &s, f3 dir) {
  if (!(s.env_integral > 0.f)) return 1.f / (4.f * PT_PI);
  f3 dn = norm3(dir);
  float th = acosf(clampf_(-dn.y, -1.f, 1.f));
  float ph = atan2f(-dn.z, dn.x) + PT_PI;
  float u = ph / (2.f * PT_PI), v = th / PT_PI;
  int x = min(max((int)(u * s.env_w), 0), s.env_w - 1);
  int y = min(max((int)(v * s.env_h), 0), s.env_h - 1);
  float st = sinf(th);
  if (!(st > 0.f)) return 0.f;
  return s.env_func[(size_t)y * s.env_w + x] / s.env_integral / (2.f * PT_PI * PT_PI * st);
}

// Hittable::pdf_value of the shape behind an area light, from o towards the
// unit direction wi (the mixture estimators): AARectangle (primitive.hpp:
// 244-256) d^2/(|cos| A) with the hit's t, Sphere (primitive.hpp:172-184)
// 1/(2 pi (1 - cos theta_max)); 0 when the ray misses the shape.
PT_D float shape_pdf(const DevLight &L, f3 o, f3 wi) {
  float t;
  if (L.type == LIGHT_QUAD) {
    if (!hit_rect(make_float4(L.k, L.d0min, L.d0max, 0.f), make_float4(L.d1min, L.d1max, (float)L.axis, 0.f), o, wi,
                  1e-4f, 1e4f, t))
      return 0.f;
    float area = (L.d0max - L.d0min) * (L.d1max - L.d1min);
    float d2 = t * t * dot3(wi, wi);
    f3 n = L.axis == 0 ? mk3(1.f, 0.f, 0.f) : (L.axis == 1 ? mk3(0.f, 1.f, 0.f) : mk3(0.f, 0.f, 1.f));
    float cosine = fabsf(dot3(wi, n) / len3(wi));
    return d2 / (cosine * area);
  }
  if (L.type == LIGHT_SPHERE) {
    if (!hit_sphere(make_float4(L.c.x, L.c.y, L.c.z, 0.f), make_float4(L.r, 0.f, 0.f, 0.f), o, wi, 1e-4f, 1e4f, t))
      return 0.f;
    f3 oc = sub3(L.c, o);
    float ctm = sqrtf(1.f - L.r * L.r / dot3(oc, oc));
    return (float)(1.0 / (2.0 * 3.14159265358979323846 * (double)(1.f - ctm)));
  }
  return 0.f;
}
// Hittable::get_sample of that shape from o (not normalised): a uniform
// point of the rect minus o (primitive.cpp:125-146), or a direction in the
// cone the sphere subtends (primitive.hpp:186-192, random.hpp:36-44).
PT_D f3 shape_sample(const DevLight &L, f3 o, float u, float v) {
  if (L.type == LIGHT_QUAD) {
    float a = L.d0min + (L.d0max - L.d0min) * u;
    float b = L.d1min + (L.d1max - L.d1min) * v;
    f3 pt = L.axis == 0 ? mk3(L.k, a, b) : (L.axis == 1 ? mk3(a, L.k, b) : mk3(a, b, L.k));
    return sub3(pt, o);
  }
  f3 dir = sub3(L.c, o);
  float d2 = dot3(dir, dir);
  float z = 1.f + v * (sqrtf(1.f - L.r * L.r / d2) - 1.f);
  float phi = 2.f * PT_PI * u;
  float x = cosf(phi) * sqrtf(1.f - z * z);
  float y = sinf(phi) * sqrtf(1.f - z * z);
  Frame b = make_frame(norm3(dir));
  return to_world(b, mk3(x, y, z));
}

// Light::get_sample + get_normal_at (light.hpp:131-147, 254-271;
// envLight.hpp:20-37,51-54)
PT_D f3 light_point(const DevLight &L, float u, float v, f3 &nrm) {
  if (L.type == LIGHT_QUAD) {
    nrm = L.n;
    return add3(add3(L.v0, scl3(L.e1, u)), scl3(L.e2, v));
  }
  if (L.type == LIGHT_SPHERE) {
    float th = 2.f * PT_PI * u;
    float ph = acosf(2.f * v - 1.f);
    f3 dir = mk3(sinf(ph) * cosf(th), sinf(ph) * sinf(th), cosf(ph));
    f3 p = add3(L.c, scl3(dir, L.r));
    nrm = norm3(sub3(p, L.c));
    return p;
  }
  f3 p = scl3(uniform_sphere(u, v), 1e5f);
  nrm = neg3(norm3(p));
  return p;
}
// Solid-angle pdf of the MIS estimator's light sampling for point p of
// light L seen from o (area lights: uniform area sampling).
PT_D float light_pdf_sa(const DevScene &s, const DevLight &L, f3 o, f3 p) {
  if (L.type == LIGHT_ENV) return env_pdf(s, sub3(p, o));
  f3 w = sub3(p, o);
  float d2 = dot3(w, w);
  f3 nl = (L.type == LIGHT_QUAD) ? L.n : norm3(sub3(p, L.c));
  float c = fabsf(dot3(nl, norm3(w)));
  if (!(c > 0.f) || !(L.area > 0.f)) return 0.f;
  return d2 / (L.area * c);
}

// ---------------------------------------------------------------------
// PerspectiveCamera::generateRay (camera.cpp:42-56): lens point uniform on
// the square [-r, r]^2 (r = focalDist/aperture/2), direction towards the
// image-plane point; a shutter time moves the camera (motion blur).
// ---------------------------------------------------------------------
PT_D void camera_ray(const DevCamera &c, float u, float v, float r0, float r1, float rt, f3 &o, f3 &d) {
  float lu = -c.radius + (c.radius - -c.radius) * r0;
  float lv = -c.radius + (c.radius - -c.radius) * r1;
  f3 pos = c.pos;
  if (c.shutter > 0.f) pos = add3(pos, scl3(sub3(c.pos1, pos), rt * c.shutter));
  f3 bl = sub3(sub3(pos, divs3(c.horiz, 2.f)), divs3(c.vert, 2.f));
  bl = sub3(bl, scl3(c.lookat, c.focal_dist));
  f3 ps = add3(pos, add3(scl3(c.h, lu), scl3(c.v, lv)));
  f3 dir = sub3(add3(add3(bl, scl3(c.horiz, u)), scl3(c.vert, v)), ps);
  if (c.normalize_dir) dir = norm3_rn(dir);
  o = ps;
  d = dir;
}

// Wave-aggregated queue append: one atomic per wave.  Must be reached by
// every lane of the wave.
PT_D uint32_t wave_append(uint32_t *counter, bool pred) {
  unsigned long long mask = __ballot(pred);
  uint32_t lane = __lane_id();
  uint32_t base = 0;
  if (mask) {
    uint32_t leader = (uint32_t)__ffsll((long long)mask) - 1u;
    uint32_t total = (uint32_t)__popcll(mask);
    if (lane == leader) base = atomicAdd(counter, total);
    base = __shfl(base, (int)leader);
  }
  return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

PT_D void wave_add_u64(unsigned long long *ctr, uint32_t v) {
  // wave reduction then one atomic
  unsigned long long x = v;
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  if (__lane_id() == 0 && x) atomicAdd(ctr, x);
}

}  // namespace pt
