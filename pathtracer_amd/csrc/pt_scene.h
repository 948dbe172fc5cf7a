// pt_scene.h -- device-resident scene and wavefront layout shared by the
// host orchestration (pt_host.cpp) and the kernels (pt_kernels.hip).
//
// HBM layout (all SoA, 256-byte aligned allocations):
//   nodes      4-wide BVH nodes, 128 B each (8 x float4): lo.x, hi.x, lo.y,
//              hi.y, lo.z, hi.z of the four children (SoA), their four child
//              codes, one unused float4.  Closest-hit traversal reads these.
//   qnodes     the same nodes at 64 B each (4 x float4, quantize_node): the
//              child boxes as 8-bit offsets, conservative.  Any-hit traversal
//              reads these.
//   prims      primitives in BVH leaf order, 48 B each (3 x float4):
//                triangle  (v0.xyz, id|type) (e1.xyz, -) (e2.xyz, -)
//                sphere    (c.xyz,  id|type) (r, -, -, -) (-)
//                rect      (k, d0min, d0max, id|type) (d1min, d1max, axis, -) (-)
//              e1 = p1-p0, e2 = p2-p0 are formed on the host exactly as
//              MeshBVHNode::hitTriangle forms them (meshBVH.cpp:93-94).
//   prim_rec   per global primitive id, 32 B (one cache line holds four):
//                geom  triangle geometric normal, sphere centre+radius,
//                      rect axis (float4)
//                info  x = material | type<<24 | flags<<28, y = light index
//                      (spheres, rects) or shading-normal record
//                      (triangles), -1 for none (uint2), then 8 B of padding
//              Shading reads both halves of a hit's record: one line.
//   tri_smooth per triangle with shading normals, 80 B: the vertex p0 and
//              edges e1, e2 (for the hit's barycentrics) and three normals.
#pragma once
#include <stdint.h>
#include "pt_math.h"
#include "pt_sampler.h"
#include "../../include/pathtracer_amd.h"

namespace pt {

enum { PRIM_TRI = 0, PRIM_SPHERE = 1, PRIM_RECT = 2, PRIM_CURVE = 3 };
// NEE+MIS with the learned guide as a second directional technique (only
// through pt_render_guided, so not in the public estimator list).
enum { PT_EST_GUIDED = 6 };
enum { LIGHT_QUAD = 1, LIGHT_SPHERE = 2, LIGHT_ENV = 3 };
enum { TEX_IMAGE = 0, TEX_CHECKER = 1 };

// A child reference: >= 0 internal node index; < 0 leaf ~(start<<4 | count-1);
// kEmptyChild marks an unused slot of a 4-wide node.
constexpr int kEmptyChild = (int)0x80000000;
#define PT_STACK4 96
// Top BVH nodes a traversal block keeps in LDS: nodes are numbered
// breadth-first, so these are the root, its 4 children and their 16 (three
// full levels of 4-wide nodes) at most.
#define PT_TOP_NODES 21
PT_HD int leaf_code(uint32_t start, uint32_t count) { return ~(int)((start << 4) | (count - 1)); }
// ---- 4-wide node, 64 bytes (4 x float4) -----------------------------------
//   [0] origin x, y, z; w: the exponent bytes e_a + 127 of the per-axis
//       scales s_a = 2^e_a (bits 0-7 x, 8-15 y, 16-23 z)
//   [1] lo_x, hi_x, lo_y, hi_y   as uint32: byte k = child k's bound
//   [2] lo_z, hi_z, 0, 0
//   [3] the four child codes (int)
// Child k's box on axis a is [fma(lo_a[k], s_a, o_a), fma(hi_a[k], s_a, o_a)],
// which quantize_node makes contain the full-precision box: the same fma
// decides at build time and in traversal, and the slab distances are monotone
// in the bounds, so a ray's box test passes wherever it passes on the full
// node.  Any-hit traversal reads these nodes: it is bound by the node bytes
// it fetches.  Closest hit keeps the 128-byte nodes, where the decode's
// arithmetic cost more than the halved fetch saved.
#define PT_QNODE_F4 4
PT_HD float pt_bits_float(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
PT_HD uint32_t pt_float_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return u;
}
// The full-precision node `in` (8 x float4: lo.x, hi.x, lo.y, hi.y, lo.z,
// hi.z of the four children, then their codes) as a quantised node `out`.
PT_HD void quantize_node(const float4 *in, float4 *out) {
  const float lo[3][4] = {{in[0].x, in[0].y, in[0].z, in[0].w}, {in[2].x, in[2].y, in[2].z, in[2].w},
                          {in[4].x, in[4].y, in[4].z, in[4].w}};
  const float hi[3][4] = {{in[1].x, in[1].y, in[1].z, in[1].w}, {in[3].x, in[3].y, in[3].z, in[3].w},
                          {in[5].x, in[5].y, in[5].z, in[5].w}};
  const int codes[4] = {(int)pt_float_bits(in[6].x), (int)pt_float_bits(in[6].y), (int)pt_float_bits(in[6].z),
                        (int)pt_float_bits(in[6].w)};
  float org[3];
  uint32_t ql[3], qh[3], eb = 0;
  for (int a = 0; a < 3; ++a) {
    float o = 0.f, m = 0.f;
    bool any = false;
    for (int k = 0; k < 4; ++k) {
      if (codes[k] == kEmptyChild) continue;
      o = any ? fminf(o, lo[a][k]) : lo[a][k];
      m = any ? fmaxf(m, hi[a][k]) : hi[a][k];
      any = true;
    }
    const float E = m - o;
    int e = -126;
    if (E > 0.f) {
      int x;
      (void)frexpf(E, &x);  // 2^(x-1) <= E < 2^x
      e = x - 9;
      if (e < -126) e = -126;
      while (e < 127 && ldexpf(255.f, e) < E) ++e;
    }
    for (;; ++e) {  // widen the scale until every child's box fits
      const float s = ldexpf(1.f, e);
      uint32_t L = 0, H = 0;
      bool ok = true;
      for (int k = 0; k < 4 && ok; ++k) {
        uint32_t l = 255u, h = 0u;  // empty slot: an inverted box
        if (codes[k] != kEmptyChild) {
          float fl = floorf((lo[a][k] - o) / s), fh = ceilf((hi[a][k] - o) / s);
          fl = fminf(fmaxf(fl, 0.f), 255.f);
          fh = fminf(fmaxf(fh, 0.f), 255.f);
          while (fl > 0.f && fmaf(fl, s, o) > lo[a][k]) fl -= 1.f;
          while (fh < 255.f && fmaf(fh, s, o) < hi[a][k]) fh += 1.f;
          ok = fmaf(fl, s, o) <= lo[a][k] && fmaf(fh, s, o) >= hi[a][k];
          l = (uint32_t)fl;
          h = (uint32_t)fh;
        }
        L |= l << (8 * k);
        H |= h << (8 * k);
      }
      if (ok || e >= 127) {
        ql[a] = L;
        qh[a] = H;
        break;
      }
    }
    org[a] = o;
    eb |= (uint32_t)(e + 127) << (8 * a);
  }
  out[0] = make_float4(org[0], org[1], org[2], pt_bits_float(eb));
  out[1] = make_float4(pt_bits_float(ql[0]), pt_bits_float(qh[0]), pt_bits_float(ql[1]), pt_bits_float(qh[1]));
  out[2] = make_float4(pt_bits_float(ql[2]), pt_bits_float(qh[2]), 0.f, 0.f);
  out[3] = in[6];
}
// float4 i of a node array in global memory, or in a block's LDS (a native
// vector type, so that the load compiles to ds_read_b128 rather than a flat
// load that waits on both memory counters).
typedef float pt_v4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(3))) pt_v4 lds_v4;
PT_HD float4 ld_f4(const float4 *p, int i) { return p[i]; }
__device__ inline float4 ld_f4(lds_v4 *p, int i) {
  const pt_v4 v = p[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
// Decodes node nd: the child boxes (SoA over the four children) and codes.
// P: const float4 * (global) or lds_v4 * (LDS).
template <class P>
PT_HD void decode_node(P nd, float (&lx)[4], float (&hx)[4], float (&ly)[4], float (&hy)[4],
                       float (&lz)[4], float (&hz)[4], int4 &ch) {
  const float4 a = ld_f4(nd, 0), b = ld_f4(nd, 1), c = ld_f4(nd, 2), d = ld_f4(nd, 3);
  const uint32_t eb = pt_float_bits(a.w);
  const float sx = pt_bits_float((eb & 255u) << 23), sy = pt_bits_float(((eb >> 8) & 255u) << 23),
              sz = pt_bits_float(((eb >> 16) & 255u) << 23);
  const uint32_t qlx = pt_float_bits(b.x), qhx = pt_float_bits(b.y), qly = pt_float_bits(b.z),
                 qhy = pt_float_bits(b.w), qlz = pt_float_bits(c.x), qhz = pt_float_bits(c.y);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lx[k] = fmaf((float)((qlx >> (8 * k)) & 255u), sx, a.x);
    hx[k] = fmaf((float)((qhx >> (8 * k)) & 255u), sx, a.x);
    ly[k] = fmaf((float)((qly >> (8 * k)) & 255u), sy, a.y);
    hy[k] = fmaf((float)((qhy >> (8 * k)) & 255u), sy, a.y);
    lz[k] = fmaf((float)((qlz >> (8 * k)) & 255u), sz, a.z);
    hz[k] = fmaf((float)((qhz >> (8 * k)) & 255u), sz, a.z);
  }
  ch = make_int4((int)pt_float_bits(d.x), (int)pt_float_bits(d.y), (int)pt_float_bits(d.z), (int)pt_float_bits(d.w));
}
PT_HD uint32_t leaf_start(int c) { return ((uint32_t)~c) >> 4; }
PT_HD uint32_t leaf_count(int c) { return (((uint32_t)~c) & 15u) + 1u; }

struct DevLight {
  int type;
  float area;
  f3 v0, e1, e2, n;  // quad: get_sample = v0 + u e1 + v e2 (light.hpp:131-138)
  f3 c;              // sphere
  float r;
  f3 le;             // texture colour * intensity
  // the AARectangle a quad light was discovered from (its shape pdf/sample)
  int axis;
  float k, d0min, d0max, d1min, d1max;
};

struct DevScene {
  const float4 *nodes;   // 128-byte nodes (closest hit)
  const float4 *qnodes;  // 64-byte quantised nodes (any hit when qany)
  uint32_t qany;         // any-hit traversal reads qnodes
  const float4 *prims;
  const float4 *prim_rec;  // 2 float4 per primitive: geom, info
  // textures (only when a material is textured, else null): texel pool,
  // 2 float4 records per texture, and the data texture coordinates need
  // (pt_host.cpp commit_textures)
  const float *tex;
  const float4 *tex_desc, *prim_tex, *tri_uv, *mesh_box;
  const float4 *tri_smooth;  // per smooth triangle: (p0,n0.x) (e1,n0.y) (e2,n0.z) (n1,n2.x) (n2.y,n2.z,-,-)
  const float4 *curve_geom;  // per curve segment: (pa, len) (axis, radius) (dP_a, -) (dP_b, -)
  const pt_material_desc *mats;
  const DevLight *lights;
  int n_lights;
  uint32_t mat_kinds;  // bit k set when some material has kind k
  int root;  // child code of the root (node 0 or a single leaf)
  uint32_t top_nodes;  // nodes [0, top_nodes) (breadth-first) are staged in LDS by traversal
  // environment (texels: (y*w + x)*3)
  const float *env;
  int env_w, env_h;
  float env_scale;
  const float *env_func, *env_cdf, *env_mcdf;
  // search ranges of the CDF inversions: bucket b of a table of G buckets
  // holds [lo, hi], the intervals the answer for any u in [b/G, (b+1)/G)
  // lies in; env_guide has W buckets per row, env_mguide H for the marginal
  const uint2 *env_guide, *env_mguide;
  float env_integral;
  int has_env;
  // the global homogeneous medium (PT_EST_VOLUME)
  float med_sa[3], med_ss[3], med_g;
  int has_medium;
  // Halton permutation tables
  const uint16_t *perm;
  const uint32_t *perm_off;
  uint32_t perm_seed;
};
PT_HD float4 prim_geom(const DevScene &s, int prim) { return s.prim_rec[2 * (size_t)prim]; }
PT_HD uint2 prim_info(const DevScene &s, int prim) {
  return *reinterpret_cast<const uint2 *>(s.prim_rec + 2 * (size_t)prim + 1);
}

struct DevCamera {
  f3 pos, pos1, h, v, horiz, vert, lookat;
  float focal_dist, radius, shutter;
  int normalize_dir;
};

// Per-render constants.
struct RenderParams {
  uint32_t width, height, spp_x, spp_y, spp, spp_begin, spp_count;
  uint32_t max_depth, sampler, estimator, rr_depth, light_samples, lx, ly, seed;
  float bg[3];
  uint64_t chunk_first;  // first path number of the chunk (pixel-major order)
  uint32_t chunk_paths;
};

// Batched rendering: many scenes rendered in one wavefront.  The global path
// space is the concatenation of the instances' pixel-major path ranges; it is
// cut into chunks of whole pixels, and a chunk into segments, one per
// instance it touches.  All of it lives in device memory, so one captured
// sequence of launches serves every chunk (the chunk index is a device
// counter that k_next_chunk advances).
struct ChunkSeg {
  uint32_t inst;        // instance
  uint32_t path_first;  // first chunk slot of the segment
  uint32_t pix_offset;  // first chunk pixel of the segment
  uint32_t pix_first;   // the instance's pixel index of that pixel
  uint32_t npix;
};
struct ChunkInfo {
  uint32_t seg_begin, n_segs;  // into BatchView::segs
  uint32_t paths, npix;        // slots and pixels of the chunk
};
struct BatchView {
  const DevScene *scenes;      // per instance
  const RenderParams *params;  // per instance (the chunk fields are unused)
  const DevCamera *cams;
  const float *offs;           // per instance 2 * spp pixel offsets, at off_first[inst]
  const uint32_t *off_first;
  float *const *films;         // per instance width * height * 3
  const ChunkSeg *segs;
  const ChunkInfo *chunks;
  uint32_t *cur;               // index of the chunk being rendered
  uint32_t *slot_inst;         // per slot: instance
  uint32_t *slot_pix;          // per slot: the instance's pixel index
  uint32_t *slot_k;            // per slot: sample index within the pixel
};

// One compacted buffer of live paths: entry i is the i-th live path of the
// current vertex, so every kernel reads and writes it with unit stride.
struct PathState {
  float *ox, *oy, *oz, *dx, *dy, *dz;  // ray
  float *t;
  int *prim;
  float *bx, *by, *bz;  // throughput
  float *ppdf;          // pdf of the BSDF sample that led here (MIS)
  float *px, *py, *pz;  // previous vertex (MIS)
  uint32_t *flags;      // bit0: previous vertex specular (or camera); bit1: reservoir pending
  uint32_t *slot;       // the path's slot: its sample and its radiance accumulator
  float *pbx, *pby, *pbz;  // throughput of the vertex whose resampled direct light is pending
};

// Wavefront buffers for `cap` path slots.  Live paths move between the two
// PathState buffers each vertex (shading appends survivors to the other
// one); radiance stays per slot, where shading, shadow rays and the film
// meet.
// Work is split into PT_PARTS partitions, each with its own counters on
// their own 128-byte lines.  Block b of every kernel serves partition
// b % PT_PARTS (grids are multiples of PT_PARTS).  Live paths stay in their
// partition for the whole chunk: shading appends the survivors of partition
// p to partition p of the other buffer, and its shadow rays to partition p
// of the shadow queue.  One counter per queue serialised ~5M atomics per
// frame on one L2 line on the 1.3M-triangle bench frame (the shadow queue's
// alone cost 23 of 63 ms of shading).
#define PT_PARTS 32
#define PT_PART_RUN 64        // slots are dealt to partitions in runs of this many
#define PT_COUNTER_STRIDE 32  // uint32 words between counters (128 bytes)
struct Wave {
  PathState st[2];  // partition p of buffer b: entries [p * part_cap, p * part_cap + live(b, p))
  float *lx, *ly, *lz;  // radiance, per slot
  uint32_t *counts;     // live(b, p), then shadow(p); see the accessors below
  unsigned long long *stats;   // CNT_* traversal / path counters
  // shadow queue: partition p holds entries [p * shadow_part_cap, p * shadow_part_cap + shadow(p))
  float *sox, *soy, *soz, *sdx, *sdy, *sdz, *stmax, *scr, *scg, *scb;
  uint32_t *sslot;
  // reservoir direct light (allocated for that estimator only): candidate i of
  // slot q at rc[(3 i + c) cap + q], its visibility at rvis[4 q + i]
  float *rc;
  uint8_t *rvis;
  // path guiding (allocated for guided renders only): the processed mixture
  // of live entry e at gmix[48 e] (with gpix, the camera vertex's mixture of
  // pixel p at gmix[48 (p - the chunk's first pixel)]); training records of
  // vertex v of slot r * grec_stride at grec[24 (r * grec_verts + v)]; the
  // pass number tags the records this pass wrote
  float *gmix;
  float *grec;
  uint32_t grec_stride, grec_verts, gpass;
  uint32_t gverts;  // vertices 0 .. gverts-1 are guided
  uint32_t gpix;    // 1: one network query per pixel at the camera vertex
  uint32_t gkinds;  // material kinds guided (g_guidable)
  float glo[3], ginv[3];  // the network's position normalisation
  uint32_t cap, part_cap, shadow_part_cap;
  uint32_t sort_keys;  // 1: shading orders each tile by material kind
  PT_HD uint32_t *live(uint32_t b, uint32_t p) const { return counts + (b * PT_PARTS + p) * PT_COUNTER_STRIDE; }
  PT_HD uint32_t *shadow(uint32_t p) const { return counts + (2 * PT_PARTS + p) * PT_COUNTER_STRIDE; }
  // work queue of partition p (k = 0: closest hit, 1: any hit, 2: shading):
  // the next unclaimed entry.  k_generate rewinds the closest-hit counters,
  // k_shade both traversal kinds and k_extend the shading ones, so every
  // launch starts from zero.
  PT_HD uint32_t *work(uint32_t k, uint32_t p) const { return counts + ((3 + k) * PT_PARTS + p) * PT_COUNTER_STRIDE; }
};
#define PT_COUNTER_WORDS (6 * PT_PARTS * PT_COUNTER_STRIDE)

enum {
  CNT_Q0 = 0, CNT_Q1 = 1, CNT_SEGMENTS = 3, CNT_SHADOW_RAYS = 4,
  CNT_NODES = 5, CNT_PRIMS = 6, CNT_SNODES = 7, CNT_SPRIMS = 8, CNT_NAN = 9, CNT_N = 16
};

}  // namespace pt
