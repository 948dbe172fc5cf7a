// pt_kernels.hip -- wavefront path-tracing kernels for gfx950.
//
// One render is a sequence of chunks of path samples (pixel-major order,
// whole pixels per chunk).  Per chunk:
//   k_generate                       camera rays for every slot
//   per vertex d:  k_extend          closest hit of the live paths
//                  k_shade<EST, MS>  material, emission, NEE shadow rays;
//                                    survivors are appended (compacted) to
//                                    the other live-path buffer
//                  k_shadow          any-hit of the shadow queue, adds the
//                                    unoccluded contributions
//   k_film_lanes / k_film_batch      per-pixel sums (fixed order)
// Live paths are stored densely, so extension and shading read and write
// them with unit stride; only radiance is addressed by slot.
// The estimator is getRayColor of integrator/maxdepth_nee.hpp:9-256
// unrolled into a loop (EST 0), or NEE+MIS (EST 1).
#include "pt_device.h"
#include "pt_guide.h"
#include "pt_launch.h"

namespace pt {

__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// ------------------------------------------------------------------ generate
// Camera rays for every slot of the chunk.  Slots are dealt to the live-path
// partitions in runs of R = 64 (one wave's worth, a pixel's samples at 64
// spp): partition p takes runs p, p + P, p + 2P, ..., so every partition holds
// an even mix of the chunk's pixels while each wave starts coherent.  Thread
// t writes entry t % m of partition t / m (m = R * ceil(n / (R P))), so the
// writes are unit-stride.  Also zeroes the radiance and sets the counters.
// Batched: the chunk comes from the batch tables; the slot's segment gives
// its instance, pixel and sample.
PT_D uint32_t find_seg(const BatchView &b, const ChunkInfo &ci, uint32_t slot) {
  uint32_t lo = ci.seg_begin, hi = ci.seg_begin + ci.n_segs - 1u;
  while (lo < hi) {  // last segment with path_first <= slot
    uint32_t mid = (lo + hi + 1u) / 2u;
    if (b.segs[mid].path_first <= slot) lo = mid;
    else hi = mid - 1u;
  }
  return lo;
}
PT_D uint32_t find_seg_pix(const BatchView &b, const ChunkInfo &ci, uint32_t pl) {
  uint32_t lo = ci.seg_begin, hi = ci.seg_begin + ci.n_segs - 1u;
  while (lo < hi) {  // last segment with pix_offset <= pl
    uint32_t mid = (lo + hi + 1u) / 2u;
    if (b.segs[mid].pix_offset <= pl) lo = mid;
    else hi = mid - 1u;
  }
  return lo;
}
template <bool BATCH>
__global__ void __launch_bounds__(256) k_generate(DevScene s0, DevCamera cam0, RenderParams rp0, Wave w,
                                                  const float *offx0, const float *offy0, BatchView b) {
  const uint32_t t = gtid();
  const uint32_t n = BATCH ? b.chunks[*b.cur].paths : rp0.chunk_paths;
  const uint32_t per = (n + PT_PART_RUN * PT_PARTS - 1) / (PT_PART_RUN * PT_PARTS) * PT_PART_RUN;
  if (t < n) {
    w.lx[t] = 0.f; w.ly[t] = 0.f; w.lz[t] = 0.f;
  }
  if (t >= per * PT_PARTS) return;
  const uint32_t part = t / per, l = t % per;
  if (l == 0) {
    const uint32_t full = n / (PT_PART_RUN * PT_PARTS), rest = n - full * PT_PART_RUN * PT_PARTS;
    const uint32_t extra = rest > part * PT_PART_RUN ? min(rest - part * PT_PART_RUN, (uint32_t)PT_PART_RUN) : 0u;
    *w.live(0, part) = full * PT_PART_RUN + extra;
    *w.live(1, part) = 0u;
    *w.work(0, part) = 0u;  // the first extension launch's queue
  }
  const uint32_t slot = ((l / PT_PART_RUN) * PT_PARTS + part) * PT_PART_RUN + l % PT_PART_RUN;
  if (slot >= n) return;
  const uint32_t e = part * w.part_cap + l;
  uint32_t pixel, k, inst = 0u;
  if (BATCH) {
    const ChunkInfo ci = b.chunks[*b.cur];
    const ChunkSeg sg = b.segs[find_seg(b, ci, slot)];
    inst = sg.inst;
    const uint32_t local = slot - sg.path_first, sppc = b.params[inst].spp_count;
    pixel = sg.pix_first + local / sppc;
    k = b.params[inst].spp_begin + local % sppc;
    b.slot_inst[slot] = inst;
    b.slot_pix[slot] = pixel;
    b.slot_k[slot] = k;
  } else {
    uint64_t q = rp0.chunk_first + slot;
    pixel = (uint32_t)(q / rp0.spp_count);
    k = rp0.spp_begin + (uint32_t)(q % rp0.spp_count);
  }
  const RenderParams &rp = BATCH ? b.params[inst] : rp0;
  const DevCamera &cam = BATCH ? b.cams[inst] : cam0;
  const float *offx = BATCH ? b.offs + b.off_first[inst] : offx0;
  const float *offy = BATCH ? offx + rp.spp : offy0;
  uint32_t pi = pixel % rp.width, pj = pixel / rp.width;
  uint32_t g = 1u + k + rp.spp * (pj + rp.height * pi);
  float u = div_rn((float)pi + offx[k], (float)rp.width);
  float v = div_rn((float)pj + offy[k], (float)rp.height);
  f3 o, d;
  camera_ray(cam, u, v, pt_rng(g, rslot(0, SLOT_LENS_X), rp.seed), pt_rng(g, rslot(0, SLOT_LENS_Y), rp.seed),
             pt_rng(g, rslot(0, SLOT_TIME), rp.seed), o, d);
  const PathState &ps = w.st[0];
  ps.ox[e] = o.x; ps.oy[e] = o.y; ps.oz[e] = o.z;
  ps.dx[e] = d.x; ps.dy[e] = d.y; ps.dz[e] = d.z;
  ps.bx[e] = 1.f; ps.by[e] = 1.f; ps.bz[e] = 1.f;
  ps.flags[e] = 1u;
  ps.ppdf[e] = 0.f;
  ps.px[e] = o.x; ps.py[e] = o.y; ps.pz[e] = o.z;
  ps.slot[e] = slot;
}

// ------------------------------------------------------------------ extend
// Traversal-stack entries kept in LDS per lane (the rest in private memory):
// 8 (code, distance) pairs for closest hit, 16 codes for any hit -- 64 bytes
// per lane, 16 KB per block, which leaves occupancy to the register budget.
// Measured on the 1.3M-triangle glass scene against the private stack alone
// (K = 0): closest hit 9% and any hit 4% faster.
#ifndef PT_K_EXT
#define PT_K_EXT 8
#endif
#ifndef PT_K_SHADOW
#define PT_K_SHADOW 16
#endif

#ifndef PT_TRAV_WAVES
#define PT_TRAV_WAVES 8
#endif
// Idle lanes a wave waits for before refilling them (64: only when all are).
// Swept on the 1.3M-triangle glass frame with the work queues: closest hit
// {16, 32, 48} -> {26.2, 25.7, 25.9} ms, any hit {2, 8, 16} -> {16.9, 15.7,
// 15.2} ms (the Disney + hair grid: any hit 12.0 -> 11.0 ms at 16); the grid
// is one resident round (8 blocks per CU).
#ifndef PT_REFILL_EXT
#define PT_REFILL_EXT 32
#endif
#ifndef PT_REFILL_SHADOW
#define PT_REFILL_SHADOW 16
#endif
// Traversal work sharing: eighths of a partition's entries split into fixed
// per-wave shares, and the entries a wave claims at a time from the rest.
#ifndef PT_STATIC_EXT
#define PT_STATIC_EXT 0
#endif
#ifndef PT_STATIC_SHADOW
#define PT_STATIC_SHADOW 0
#endif
// Shading: tiles claimed through a counter (1) or dealt round robin (0).
#ifndef PT_SHADE_QUEUE
#define PT_SHADE_QUEUE 1
#endif
#ifndef PT_SHADE_QUEUE_MIN
#define PT_SHADE_QUEUE_MIN 4
#endif
#ifndef PT_CHUNK_EXT
#define PT_CHUNK_EXT 64
#endif
#ifndef PT_CHUNK_SHADOW
#define PT_CHUNK_SHADOW 64
#endif
template <bool COUNT, bool BATCH>
__global__ void __launch_bounds__(PT_BLOCK) __attribute__((amdgpu_waves_per_eu(PT_TRAV_WAVES))) k_extend(DevScene s, Wave w, int qi, BatchView b) {
  __shared__ int lds[PT_TRAV_LDS(PT_K_EXT, true)];
  __shared__ float4 top[8 * PT_TOP_NODES];
  stage_top_nodes<false>(s, top, BATCH);
  const uint32_t part = blockIdx.x % PT_PARTS, nb = gridDim.x / PT_PARTS;
  const uint32_t n = *w.live(qi, part), base = part * w.part_cap;
  const PathState &ps = w.st[qi];
  // the shading that follows appends to the other buffer and the shadow queue
  if (blockIdx.x == 0 && threadIdx.x < PT_PARTS) {
    *w.live(qi ^ 1, threadIdx.x) = 0u;
    *w.shadow(threadIdx.x) = 0u;
    *w.work(2, threadIdx.x) = 0u;  // the shading launch's tile queue
  }
  uint32_t nn = 0, np = 0, segs = 0;
  const uint32_t wpb = blockDim.x / 64u, wid = (blockIdx.x / PT_PARTS) * wpb + threadIdx.x / 64u;
  trace_queue<false, false, COUNT, PT_K_EXT, PT_REFILL_EXT, PT_STATIC_EXT, PT_CHUNK_EXT, BATCH>(
      s, lds, top, n, w.work(0, part), wid, nb * wpb,
      [&](uint32_t i, f3 &o, f3 &d, float &tmin, float &tmax, const DevScene *&sc) {
        i += base;
        if (BATCH) sc = &b.scenes[b.slot_inst[ps.slot[i]]];
        o = mk3(ps.ox[i], ps.oy[i], ps.oz[i]);
        d = mk3(ps.dx[i], ps.dy[i], ps.dz[i]);
        tmin = 1e-4f;
        tmax = 1e5f;
      },
      [&](uint32_t i, bool hit, float t, int prim) {
        i += base;
        ps.t[i] = t;
        ps.prim[i] = hit ? prim : -1;
        ++segs;
      },
      nn, np);
  if (COUNT) {
    wave_add_u64(&w.stats[CNT_NODES], nn);
    wave_add_u64(&w.stats[CNT_PRIMS], np);
    wave_add_u64(&w.stats[CNT_SEGMENTS], segs);
  }
}

// ------------------------------------------------------------------ shadow
// VIS: the entries are reservoir candidates (slot field = 4 slot + candidate),
// whose visibility is recorded instead of adding a contribution.
template <bool COUNT, bool VIS, bool BATCH, bool QN>
__global__ void __launch_bounds__(PT_BLOCK) __attribute__((amdgpu_waves_per_eu(PT_TRAV_WAVES))) k_shadow(DevScene s, Wave w, BatchView b) {
  __shared__ int lds[PT_TRAV_LDS(PT_K_SHADOW, false)];
  __shared__ float4 top[node_f4<QN>() * PT_TOP_NODES];
  stage_top_nodes<QN>(s, top, BATCH);
  // blocks b, b + P, b + 2P, ... share partition p = b % P (the grid is a multiple of P)
  const uint32_t p = blockIdx.x % PT_PARTS, nb = gridDim.x / PT_PARTS;
  const uint32_t n = min(*w.shadow(p), w.shadow_part_cap);
  const uint32_t base = p * w.shadow_part_cap;
  uint32_t nn = 0, np = 0, rays = 0;
  const uint32_t wpb = blockDim.x / 64u, wid = (blockIdx.x / PT_PARTS) * wpb + threadIdx.x / 64u;
  trace_queue<true, QN, COUNT, PT_K_SHADOW, PT_REFILL_SHADOW, PT_STATIC_SHADOW, PT_CHUNK_SHADOW, BATCH>(
      s, lds, top, n, w.work(1, p), wid, nb * wpb,
      [&](uint32_t i, f3 &o, f3 &d, float &tmin, float &tmax, const DevScene *&sc) {
        i += base;
        if (BATCH) sc = &b.scenes[b.slot_inst[VIS ? w.sslot[i] / 4u : w.sslot[i]]];
        o = mk3(w.sox[i], w.soy[i], w.soz[i]);
        d = mk3(w.sdx[i], w.sdy[i], w.sdz[i]);
        tmin = 0.001f;
        tmax = w.stmax[i];
      },
      [&](uint32_t i, bool occ, float, int) {
        i += base;
        ++rays;
        if (!occ) {
          uint32_t slot = w.sslot[i];
          if (VIS) {
            w.rvis[slot] = 1u;
          } else {
            atomicAdd(&w.lx[slot], w.scr[i]);
            atomicAdd(&w.ly[slot], w.scg[i]);
            atomicAdd(&w.lz[slot], w.scb[i]);
          }
        }
      },
      nn, np);
  if (COUNT) {
    wave_add_u64(&w.stats[CNT_SNODES], nn);
    wave_add_u64(&w.stats[CNT_SPRIMS], np);
    wave_add_u64(&w.stats[CNT_SHADOW_RAYS], rays);
  }
}

// ------------------------------------------------------------------ shade
// Appends one shadow ray to this block's partition of the shadow queue.
// Called from divergent code: the compiler turns the uniform-address atomic
// into one atomic per wave (lane-ordered offsets).  A partition holds 4
// entries per live-path entry of the same partition, so the capacity check
// never drops a ray.
PT_D void emit_shadow(const Wave &w, f3 p, f3 d, float tmax, f3 c, uint32_t slot) {
  const uint32_t part = blockIdx.x % PT_PARTS;
  uint32_t idx = atomicAdd(w.shadow(part), 1u);
  if (idx < w.shadow_part_cap) {
    idx += part * w.shadow_part_cap;
    w.sox[idx] = p.x; w.soy[idx] = p.y; w.soz[idx] = p.z;
    w.sdx[idx] = d.x; w.sdy[idx] = d.y; w.sdz[idx] = d.z;
    w.stmax[idx] = tmax;
    w.scr[idx] = c.x; w.scg[idx] = c.y; w.scb[idx] = c.z;
    w.sslot[idx] = slot;
  }
}

PT_D f3 miss_radiance(const DevScene &s, const RenderParams &rp, f3 d) {
  if (s.has_env) return env_lookup(s, d);
  return mk3(rp.bg[0], rp.bg[1], rp.bg[2]);
}
PT_D float power_w(float a, float b) { return (a * a) / (a * a + b * b); }

// ------------------------------------------------------------------ material-key order
// Shading order of one 256-entry tile: a stable counting sort of the entries
// by a 4-bit key (the hit's material kind; 0 = miss, 15 = no entry), so that
// each wave shades one kind where the tile allows.  Entries stay in their
// tile, so the permuted state loads touch the same cache lines as the
// unpermuted ones.  Returns the tile index of the entry this thread shades.
// Every thread of the block calls it (it synchronises the block).
PT_D uint32_t tile_order16(uint32_t key, uint32_t (*cnt)[16], uint32_t *first, uint16_t *perm) {
  const uint32_t lane = __lane_id(), wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t rank = 0;
#pragma unroll
  for (uint32_t k = 0; k < 16u; ++k) {
    const unsigned long long m = __ballot(key == k);
    if (key == k) rank = (uint32_t)__popcll(m & below);
    if (lane == 0) cnt[wv][k] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t k = 0; k < 16u; ++k) {
      first[k] = acc;
      for (uint32_t v = 0; v < nw; ++v) acc += cnt[v][k];
    }
  }
  __syncthreads();
  uint32_t pos = first[key] + rank;
  for (uint32_t v = 0; v < wv; ++v) pos += cnt[v][key];
  perm[pos] = (uint16_t)threadIdx.x;
  __syncthreads();
  const uint32_t src = perm[threadIdx.x];
  __syncthreads();  // the next tile reuses the tables
  return src;
}

// EST 0: getRayColor (maxdepth_nee.hpp:9-256).  EST 1: NEE + MIS.
// Occupancy target of the shading kernel (waves per SIMD): the register
// allocator spills the rarely used material state to reach it.  On the
// 1.3M-triangle glass frame, shading took 26.3 / 26.2 / 24.8 / 24.4 ms at
// 3 / 4 / 5 / 6 waves (whole frame 83.9 / 83.2-83.8 / 83.0 / 82.6 ms).
#ifndef PT_SHADE_WAVES
#define PT_SHADE_WAVES 6
#endif
// The same for the all-kinds variant (Disney and hair): 3 waves/SIMD (150
// VGPRs, no spills) shaded the 64 Disney + 8 hair spheres in 116.5 ms against
// 109.7 ms at 4 waves (128 VGPRs, 48 bytes per lane spilled).
#ifndef PT_SHADE_WAVES_ALL
#define PT_SHADE_WAVES_ALL 4
#endif
// The pixel, sample index and random-stream index of a path slot.
template <bool BATCH>
PT_D void path_index(const RenderParams &rp, const BatchView &b, uint32_t slot, uint32_t &pixel, uint32_t &k,
                     uint32_t &g) {
  if (BATCH) {
    pixel = b.slot_pix[slot];
    k = b.slot_k[slot];
  } else {
    const uint64_t q = rp.chunk_first + slot;
    pixel = (uint32_t)(q / rp.spp_count);
    k = rp.spp_begin + (uint32_t)(q % rp.spp_count);
  }
  const uint32_t pi = pixel % rp.width, pj = pixel / rp.width;
  g = 1u + k + rp.spp * (pj + rp.height * pi);
}

// A path whose extension ray left the scene, under the reference estimator,
// NEE+MIS or the guided estimator: the environment or background radiance it
// adds (weighted against the previous vertex's light sample but for the
// reference estimator), and the path ends; guided, the previous vertex's
// training record is completed as k_shade completes it.  k_shade's own miss
// case, for the entries its tiles shade as they read them.
template <int EST, bool BATCH>
PT_D void shade_miss(const DevScene &s0, const RenderParams &rp0, const Wave &w, const PathState &in, uint32_t i,
                     uint32_t d, const BatchView &b) {
  const uint32_t slot = in.slot[i];
  const uint32_t inst = BATCH ? b.slot_inst[slot] : 0u;
  const DevScene &s = BATCH ? b.scenes[inst] : s0;
  const RenderParams &rp = BATCH ? b.params[inst] : rp0;
  const f3 dir = mk3(in.dx[i], in.dy[i], in.dz[i]);
  const f3 beta = mk3(in.bx[i], in.by[i], in.bz[i]);
  const bool prev_spec = in.flags[i] & 1u;
  float *prec = nullptr;
  if constexpr (EST == PT_EST_GUIDED) {
    if (w.grec && d >= 1u && d - 1u < w.grec_verts && slot % w.grec_stride == 0u) {
      prec = w.grec + ((size_t)(slot / w.grec_stride) * w.grec_verts + (d - 1u)) * PT_G_REC;
      prec[13] = w.lx[slot];
      prec[14] = w.ly[slot];
      prec[15] = w.lz[slot];
      prec[17] = __uint_as_float(w.gpass);
      prec[18] = prec[19] = prec[20] = 0.f;
    }
  }
  f3 L = s3(0.f);
  if (EST == PT_EST_REFERENCE_NEE) {
    L = mul3(beta, miss_radiance(s, rp, dir));
  } else if (prev_spec) {
    if (d < rp.max_depth) L = mul3(beta, miss_radiance(s, rp, dir));
  } else if (s.has_env) {
    const float pl = env_pdf(s, dir) / (float)s.n_lights;
    const f3 le = mul3(beta, env_lookup(s, dir));
    const float wm = power_w(in.ppdf[i], pl);
    L = scl3(le, wm);
    if (EST == PT_EST_GUIDED && prec) {  // the target is the unweighted radiance
      prec[18] = le.x * (1.f - wm); prec[19] = le.y * (1.f - wm); prec[20] = le.z * (1.f - wm);
    }
  } else {
    L = mul3(beta, miss_radiance(s, rp, dir));
  }
  if (L.x != 0.f || L.y != 0.f || L.z != 0.f || !finite3(L)) {
    w.lx[slot] += L.x;
    w.ly[slot] += L.y;
    w.lz[slot] += L.z;
  }
}

template <int EST, uint32_t MS, bool BATCH>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MS == 0xffffffffu ? PT_SHADE_WAVES_ALL : PT_SHADE_WAVES))) k_shade(DevScene s0, RenderParams rp0, Wave w, int qi, uint32_t d, BatchView b) {
  const uint32_t part = blockIdx.x % PT_PARTS, nb = gridDim.x / PT_PARTS;
  const uint32_t n = *w.live(qi, part), base = part * w.part_cap;
  const PathState &in = w.st[qi];
  const PathState &out = w.st[qi ^ 1];
  uint32_t *cout = w.live(qi ^ 1, part);
  // the traversal queues of the launches that follow (the shadow rays this
  // vertex emits, the next extension) start from zero; the ones before ended
  if (blockIdx.x == 0 && threadIdx.x < 2 * PT_PARTS) *w.work(threadIdx.x / PT_PARTS, threadIdx.x % PT_PARTS) = 0u;
  __shared__ uint32_t o_cnt[4][16], o_first[16];
  __shared__ uint16_t o_perm[256];
  // The block reads its tiles of 256 entries into two LDS lists and shades
  // each list 256 entries at a time.  Under the reference estimator, NEE+MIS
  // and the guided estimator (without the material-key order) a miss is
  // shaded as it is read, a few instructions, and only hits join a list; hair
  // hits join the second list.  The material code then runs on full waves rather than on the hits
  // among a tile's misses, and a wave runs the hair BSDF or the others, not
  // both.  Shading of the Disney + hair grid frame: 88.6 ms with every entry
  // in place, 86.4 with misses left out, 76.4 with hair apart; the glass /
  // HDRI frame: 25.1 -> 21.4 ms.
  __shared__ uint32_t o_list[2][512];
  __shared__ uint32_t o_n[2];
  constexpr bool SPLIT = EST == PT_EST_REFERENCE_NEE || EST == PT_EST_NEE_MIS || EST == PT_EST_GUIDED;
  const bool split = SPLIT && !w.sort_keys;  // uniform
  // The partition's tiles go to its blocks through the counter work(2, part)
  // (rewound by the extension kernel before): a block claims its next tile
  // while it reads the current one, so blocks that drew cheap tiles take more.
  // Below PT_SHADE_QUEUE_MIN tiles per block the tiles are dealt round robin
  // (a launch that short would wait on the counter).  Measured against round
  // robin: shading of the Disney + hair grid frame 60.4 -> 54.5 ms, of the
  // glass / HDRI frame 20.3 -> 19.5 ms.
  __shared__ uint32_t o_tile[2];
  uint32_t *const tq = w.work(2, part);
  const bool queued = PT_SHADE_QUEUE && n >= (uint32_t)PT_SHADE_QUEUE_MIN * 256u * nb;  // uniform
  if (threadIdx.x < 2u) o_n[threadIdx.x] = 0u;
  if (queued && threadIdx.x == 0) o_tile[0] = atomicAdd(tq, 256u);
  __syncthreads();
  uint32_t t0 = queued ? o_tile[0] : (blockIdx.x / PT_PARTS) * 256u;
  for (uint32_t it = 0;; ++it) {
    if (queued && threadIdx.x == 0) o_tile[(it + 1u) & 1u] = atomicAdd(tq, 256u);
    if (t0 < n) {
      const uint32_t e = t0 + threadIdx.x;
      bool push = e < n;
      uint32_t cls = 0u;
      if (split && push) {
        const int pr = in.prim[base + e];
        if (pr < 0) {
          shade_miss<EST, BATCH>(s0, rp0, w, in, base + e, d, b);
          push = false;
        } else if (PT_IN(MS, PT_MAT_HAIR)) {
          const DevScene &s = BATCH ? b.scenes[b.slot_inst[in.slot[base + e]]] : s0;
          if (s.mat_kinds & (1u << PT_MAT_HAIR))
            cls = s.mats[prim_info(s, pr).x & 0xffffffu].kind == PT_MAT_HAIR ? 1u : 0u;
        }
      }
#pragma unroll
      for (uint32_t c = 0; c < 2u; ++c) {
        const bool pc = push && cls == c;
        const unsigned long long m = __ballot(pc);
        uint32_t wb = 0u;
        if (__lane_id() == 0u && m) wb = atomicAdd(&o_n[c], (uint32_t)__popcll(m));
        wb = __shfl(wb, 0);
        if (pc) o_list[c][wb + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull))] = e;
      }
    }
    __syncthreads();
    const uint32_t cnt0 = o_n[0], cnt1 = o_n[1];
    const uint32_t t_next = queued ? o_tile[(it + 1u) & 1u] : t0 + nb * 256u;
    const bool last = t_next >= n;  // the block's last tile (uniform)
    __syncthreads();
    for (uint32_t c = 0; c < 2u; ++c) {
      uint32_t cnt = c ? cnt1 : cnt0;
      while (cnt >= 256u || (last && cnt > 0u)) {
        uint32_t jl = threadIdx.x;
        if (w.sort_keys) {  // block-uniform
          uint32_t key = 15u;
          if (jl < cnt) {
            const uint32_t e = base + o_list[c][jl];
            const int pr = in.prim[e];
            key = 0u;
            if (pr >= 0) {
              const DevScene &s = BATCH ? b.scenes[b.slot_inst[in.slot[e]]] : s0;
              key = s.mats[prim_info(s, pr).x & 0xffffffu].kind & 15u;
            }
          }
          jl = tile_order16(key, o_cnt, o_first, o_perm);
        }
        const bool active = jl < cnt;
        const uint32_t i = base + (active ? o_list[c][jl] : 0u);
        bool cont = false;
        uint32_t slot = 0;
        f3 p = s3(0.f), wi = s3(0.f), beta = s3(0.f), pprev = s3(0.f), pbeta = s3(0.f);
        float ppdf = 0.f;
        uint32_t fl = 0u;
        if (active) {
          slot = in.slot[i];
          // the path's scene and render parameters (batched: its instance's)
          const uint32_t inst = BATCH ? b.slot_inst[slot] : 0u;
          const DevScene &s = BATCH ? b.scenes[inst] : s0;
          const RenderParams &rp = BATCH ? b.params[inst] : rp0;
          const uint32_t D = rp.max_depth;
          const int N = s.n_lights;
          uint32_t pixel, k, g;
          path_index<BATCH>(rp, b, slot, pixel, k, g);
          f3 o = mk3(in.ox[i], in.oy[i], in.oz[i]);
          f3 dir = mk3(in.dx[i], in.dy[i], in.dz[i]);
          float t = in.t[i];
          int prim = in.prim[i];
          beta = mk3(in.bx[i], in.by[i], in.bz[i]);
          fl = in.flags[i];
          ppdf = in.ppdf[i];
          pprev = mk3(in.px[i], in.py[i], in.pz[i]);
          f3 L = s3(0.f);  // radiance this vertex adds (emission); NEE goes through shadow rays
          bool prev_spec = fl & 1u;
          bool done = false;
          // guided: the training record of vertex d - 1, completed here (the slot
          // holds everything up to and including that vertex's direct light)
          float *prec = nullptr;
          if constexpr (EST == PT_EST_GUIDED) {
            if (w.grec && d >= 1u && d - 1u < w.grec_verts && slot % w.grec_stride == 0u) {
              prec = w.grec + ((size_t)(slot / w.grec_stride) * w.grec_verts + (d - 1u)) * PT_G_REC;
              prec[13] = w.lx[slot];
              prec[14] = w.ly[slot];
              prec[15] = w.lz[slot];
              prec[17] = __uint_as_float(w.gpass);
              prec[18] = prec[19] = prec[20] = 0.f;
            }
          }
          if constexpr (EST == PT_EST_RESERVOIR) {
            // getRayColor of maxdepth_reservoir_di.hpp:9-131
            const f3 bg = mk3(rp.bg[0], rp.bg[1], rp.bg[2]);
            const uint32_t ns = min(rp.light_samples, 4u);
            (void)prev_spec;
            // the previous vertex's resampled direct light, now that its candidates' visibility is known
            f3 pend = s3(0.f);
            if (fl & 2u) {
              const uint32_t dp = d - 1u;
              f3 sample = s3(0.f);
              float W = 0.f;
              for (uint32_t si = 0; si < ns; ++si) {
                const size_t c0 = (size_t)(3u * si) * w.cap + slot;
                f3 c = mk3(w.rc[c0], w.rc[c0 + w.cap], w.rc[c0 + 2 * (size_t)w.cap]);
                f3 cand = scl3(c, w.rvis[4u * slot + si] ? 1.f : 0.f);
                float wgt = len3(cand);
                W += wgt;
                if (pt_rng(g, rslot(dp, SLOT_RIS + si), rp.seed) < wgt / W) sample = cand;
              }
              pend = mul3(mk3(in.pbx[i], in.pby[i], in.pbz[i]), scl3(sample, W / (float)ns));
            }
            SurfHit h;
            const pt_material_desc *mat = nullptr;
            bool drop = false;
            if (prim >= 0) {
              h = surface_tex<MS>(s, prim, o, dir, t);
              mat = &s.mats[h.mat];
              drop = d > 0u && mat->kind == PT_MAT_EMITTER;  // :113-115: the vertex returns 0
            }
            fl = 0u;
            if (!drop) {
              L = add3(L, pend);
              if (prim < 0 || d >= D) {
                L = add3(L, mul3(beta, bg));
              } else {
                p = h.p;
                float uc = sample_dim(s, rp.sampler, rp.seed, rp.spp_x, rp.spp_y, pixel, k, g, 3u * d);
                float u0 = sample_dim(s, rp.sampler, rp.seed, rp.spp_x, rp.spp_y, pixel, k, g, 3u * d + 1u);
                float u1 = sample_dim(s, rp.sampler, rp.seed, rp.spp_x, rp.spp_y, pixel, k, g, 3u * d + 2u);
                float aux = pt_rng(g, rslot(d, SLOT_AUX), rp.seed);
                f3 wo = neg3(norm3(dir));
                Scatter sc;
                if (!bsdf_scatter<MS>(*mat, wo, h.sp, uc, u0, u1, aux, sc)) {
                  L = add3(L, mul3(beta, bsdf_emit(*mat, dir, h.sp)));
                } else {
                  float pdf;
                  if (sc.flags & BX_SPECULAR) {  // no pdf of its own: CosinePDF(n) (:39-41)
                    wi = cosine_dir(h.sp.n, u0, u1);
                    pdf = cosine_pdf(h.sp.n, wi);
                  } else {
                    wi = norm3(sc.wi);
                    pdf = bsdf_pdf<MS>(*mat, wo, wi, h.sp);
                  }
                  const float cs = fabsf(dot3(h.sp.n, wi));
                  if (N > 0 && ns > 0u) {
                    for (uint32_t si = 0; si < ns; ++si) {
                      uint32_t ii = si / rp.ly, jj = si % rp.ly;
                      float a0 = div_rn((float)ii, (float)rp.lx), a1 = div_rn((float)(ii + 1u), (float)rp.lx);
                      float b0 = div_rn((float)jj, (float)rp.ly), b1 = div_rn((float)(jj + 1u), (float)rp.ly);
                      float ox = a0 + (a1 - a0) * pt_rng(g, rslot(d, SLOT_NEE_OFF + 2u * si), rp.seed);
                      float oy = b0 + (b1 - b0) * pt_rng(g, rslot(d, SLOT_NEE_OFF + 2u * si + 1u), rp.seed);
                      int li = (int)(pt_rng(g, rslot(d, SLOT_NEE_LIGHT + si), rp.seed) * (float)N) % N;
                      const DevLight &Lg = s.lights[li];
                      f3 nl;
                      f3 lp = light_point(Lg, ox, oy, nl);
                      f3 ld = sub3(lp, p);
                      f3 lwi = norm3(ld);
                      float len = len3(ld);
                      float noi = fmaxf(dot3(h.sp.n, lwi), 0.f);
                      float nol = fmaxf(dot3(nl, neg3(lwi)), 0.f);
                      float G = nol * noi / dot3(ld, ld);
                      f3 f = bsdf_f<MS>(*mat, wo, lwi, h.sp);
                      f3 le = Lg.type == LIGHT_ENV ? env_lookup_uv(s, ox, oy) : Lg.le;
                      float area = Lg.type == LIGHT_ENV ? INFINITY : Lg.area;
                      f3 c = scl3(mul3(f, le), area * G);
                      const size_t c0 = (size_t)(3u * si) * w.cap + slot;
                      w.rc[c0] = c.x; w.rc[c0 + w.cap] = c.y; w.rc[c0 + 2 * (size_t)w.cap] = c.z;
                      w.rvis[4u * slot + si] = 0u;
                      // only the visibility of a non-zero candidate can matter
                      if (c.x != 0.f || c.y != 0.f || c.z != 0.f || !finite3(c))
                        emit_shadow(w, p, lwi, len - 0.001f, c, 4u * slot + si);
                    }
                    pbeta = beta;
                    fl = 2u;
                  }
                  beta = scl3(mul3(beta, sc.att), cs / pdf);
                  cont = d + 1u <= D;  // vertex D only decides whether this vertex counts
                }
              }
            }
          } else if constexpr (EST == PT_EST_VOLUME) {
            // getRayColor of maxdepth_volume.hpp:72-118 with estimate_l_scatter1 (:10-70):
            // single scattering in the global medium; the path ends here
            if (s.has_medium) {
              const f3 sa = mk3(s.med_sa[0], s.med_sa[1], s.med_sa[2]);
              const f3 ss = mk3(s.med_ss[0], s.med_ss[1], s.med_ss[2]);
              const f3 st = add3(sa, ss);
              const float smax = fmaxf(st.x, fmaxf(st.y, st.z));
              const float u = sample_dim(s, rp.sampler, rp.seed, rp.spp_x, rp.spp_y, pixel, k, g, 0u);
              const float tf = -logf(1.f - u) / smax;
              const float th = prim >= 0 ? t : 3.402823466e+38f;  // HitRecord::t = FLT_MAX on a miss
              if (tf < th) {
                const f3 ps = add3(o, scl3(dir, tf));
                const f3 wo = neg3(norm3(dir));
                const uint32_t ns = min(rp.light_samples, 4u);
                // albedo * L_s1 / sigma_t: the reference divides by max(sigma_t) once more than
                // the free-flight pdf asks for
                const f3 alb = mk3(ss.x / st.x, ss.y / st.y, ss.z / st.z);
                for (uint32_t si = 0; N > 0 && si < ns; ++si) {
                  const int li = (int)(pt_rng(g, rslot(0u, SLOT_NEE_LIGHT + si), rp.seed) * (float)N) % N;
                  const DevLight &Lg = s.lights[li];
                  f3 nl;
                  const f3 lp = light_point(Lg, pt_rng(g, rslot(0u, SLOT_NEE_OFF + 2u * si), rp.seed),
                                            pt_rng(g, rslot(0u, SLOT_NEE_OFF + 2u * si + 1u), rp.seed), nl);
                  const f3 dv = sub3(ps, lp);
                  const f3 dl = norm3(dv);
                  const float dist = len3(dv);
                  if (dist < 1e-3f) continue;  // gl::epsilon
                  const f3 le = (dot3(nl, dl) > 0.f) ? Lg.le : s3(0.f);
                  const float pdf = 1.f / (Lg.area * (float)N);
                  const float cosl = fabsf(dot3(nl, dl));
                  const float ph = hg_phase(dot3(wo, neg3(dl)), s.med_g);
                  const f3 T = mk3(expf(-st.x * dist), expf(-st.y * dist), expf(-st.z * dist));
                  f3 c = divs3(divs3(mul3(scl3(le, cosl * ph), T), pdf), dist * dist);
                  c = divs3(mul3(alb, divs3(c, (float)ns)), smax);
                  if ((c.x > 0.f || c.y > 0.f || c.z > 0.f) || !finite3(c))
                    emit_shadow(w, lp, dl, dist - 0.001f, c, slot);
                }
              } else if (prim >= 0) {
                const SurfHit h = surface_tex<MS>(s, prim, o, dir, t);
                const pt_material_desc &m = s.mats[h.mat];
                if (m.kind == PT_MAT_EMITTER) L = add3(L, bsdf_emit(m, dir, h.sp));
              }
            }
            cont = false;
            (void)prev_spec;
          } else if constexpr (EST == PT_EST_NAIVE || EST == PT_EST_ROULETTE) {
            // maxdepth_naive.hpp:9-75 / roulette_naive.hpp:9-78: no NEE; misses and
            // the naive depth limit return bg_color (the environment is not consulted)
            const f3 bg = mk3(rp.bg[0], rp.bg[1], rp.bg[2]);
            const int NA = N - (s.has_env ? 1 : 0);  // area lights come first
            (void)prev_spec;
            if (prim < 0) {
              L = add3(L, mul3(beta, bg));
            } else {
              SurfHit h = surface_tex<MS>(s, prim, o, dir, t);
              const pt_material_desc *mat = &s.mats[h.mat];
              p = h.p;
              float uc = sample_dim(s, rp.sampler, rp.seed, rp.spp_x, rp.spp_y, pixel, k, g, 3u * d);
              float u0 = sample_dim(s, rp.sampler, rp.seed, rp.spp_x, rp.spp_y, pixel, k, g, 3u * d + 1u);
              float u1 = sample_dim(s, rp.sampler, rp.seed, rp.spp_x, rp.spp_y, pixel, k, g, 3u * d + 2u);
              float aux = pt_rng(g, rslot(d, SLOT_AUX), rp.seed);
              f3 wo = neg3(norm3(dir));
              Scatter sc;
              if (!bsdf_scatter<MS>(*mat, wo, h.sp, uc, u0, u1, aux, sc)) {
                L = add3(L, mul3(beta, bsdf_emit(*mat, dir, h.sp)));
              } else {
                float q = 1.f;
                bool alive = true;
                if (EST == PT_EST_ROULETTE) {
                  q = fminf(fmaxf(maxc3(sc.att), 0.5f), 1.f);
                  alive = !(pt_rng(g, rslot(d, SLOT_RR), rp.seed) > q);
                }
                if (alive) {
                  if (sc.flags & BX_SPECULAR) {
                    beta = mul3(beta, divs3(sc.att, q));
                    wi = sc.wi;
                  } else {
                    // MixedPDF of light_samples HittablePDFs and the material's pdf
                    int li[4] = {0, 0, 0, 0};
                    const uint32_t nl = NA > 0 ? min(rp.light_samples, 4u) : 0u;
                    for (uint32_t si = 0; si < nl; ++si)
                      li[si] = (int)(pt_rng(g, rslot(d, SLOT_NEE_LIGHT + si), rp.seed) * (float)NA) % NA;
                    const uint32_t ncomp = nl + 1u;
                    const uint32_t comp =
                        (uint32_t)floorf(pt_rng(g, rslot(d, SLOT_MIX), rp.seed) * (float)ncomp) % ncomp;
                    if (comp < nl) {
                      float e0 = pt_rng(g, rslot(d, SLOT_NEE_OFF), rp.seed);
                      float e1 = pt_rng(g, rslot(d, SLOT_NEE_OFF + 1u), rp.seed);
                      int lc = li[0];
                      for (uint32_t si = 1; si < nl; ++si)
                        if (si == comp) lc = li[si];
                      wi = norm3(shape_sample(s.lights[lc], p, e0, e1));
                    } else {
                      wi = norm3(sc.wi);
                    }
                    const float wgt = 1.f / (float)ncomp;
                    float pdf = 0.f;
                    for (uint32_t si = 0; si < nl; ++si) pdf += wgt * shape_pdf(s.lights[li[si]], p, wi);
                    pdf += wgt * bsdf_pdf<MS>(*mat, wo, wi, h.sp);
                    f3 f = divs3(bsdf_f<MS>(*mat, wo, wi, h.sp), q);
                    float cs = fabsf(dot3(h.sp.n, wi));
                    beta = scl3(mul3(beta, f), cs / pdf);
                  }
                  if (d + 1u < D) {
                    cont = beta.x != 0.f || beta.y != 0.f || beta.z != 0.f;
                  } else if (EST == PT_EST_NAIVE) {
                    L = add3(L, mul3(beta, bg));  // getRayColor(.., 0) returns bg_color
                  }
                  fl = 0u;
                }
              }
            }
          } else {
            if (prim < 0) {
              if (EST == 0) {
                L = add3(L, mul3(beta, miss_radiance(s, rp, dir)));
              } else if (prev_spec) {
                if (d < D) L = add3(L, mul3(beta, miss_radiance(s, rp, dir)));
              } else if (s.has_env) {
                float pl = env_pdf(s, dir) / (float)N;
                const f3 le = mul3(beta, env_lookup(s, dir));
                const float wm = power_w(ppdf, pl);
                L = add3(L, scl3(le, wm));
                if (EST == PT_EST_GUIDED && prec) {  // the target is the unweighted radiance
                  prec[18] = le.x * (1.f - wm); prec[19] = le.y * (1.f - wm); prec[20] = le.z * (1.f - wm);
                }
              } else {
                L = add3(L, mul3(beta, miss_radiance(s, rp, dir)));
              }
              done = true;
            }
            SurfHit h;
            const pt_material_desc *mat = nullptr;
            if (!done) {
              h = surface_tex<MS>(s, prim, o, dir, t);
              mat = &s.mats[h.mat];
              if (mat->kind == PT_MAT_EMITTER) {
                if (EST == 0) {
                  // A non-specular predecessor excludes emitter hits (maxdepth_nee.hpp:246-247);
                  // camera and specular chains take DiffuseEmitter::emit.
                  if (prev_spec) L = add3(L, mul3(beta, bsdf_emit(*mat, dir, h.sp)));
                } else if (prev_spec || h.light < 0) {
                  if (d < D) L = add3(L, mul3(beta, bsdf_emit(*mat, dir, h.sp)));
                } else {
                  const DevLight &Lg = s.lights[h.light];
                  // one-sided QuadLight/SphereLight::L_emit (light.hpp:157-165, 296-303): the quad's
                  // normal follows its vertex order, the sphere's points outwards
                  f3 nl = (Lg.type == LIGHT_QUAD) ? Lg.n : h.ng;
                  f3 le = (dot3(nl, neg3(norm3(dir))) > 0.f) ? Lg.le : s3(0.f);
                  float pl = light_pdf_sa(s, Lg, pprev, h.p) / (float)N;
                  const float wm = power_w(ppdf, pl);
                  L = add3(L, scl3(mul3(beta, le), wm));
                  if (EST == PT_EST_GUIDED && prec) {  // the target is the unweighted radiance
                    const f3 r = scl3(mul3(beta, le), 1.f - wm);
                    prec[18] = r.x; prec[19] = r.y; prec[20] = r.z;
                  }
                }
                done = true;
              }
            }
            if (!done && d >= D) done = true;
            if (!done) {
              p = h.p;
              float uc = sample_dim(s, rp.sampler, rp.seed, rp.spp_x, rp.spp_y, pixel, k, g, 3u * d);
              float u0 = sample_dim(s, rp.sampler, rp.seed, rp.spp_x, rp.spp_y, pixel, k, g, 3u * d + 1u);
              float u1 = sample_dim(s, rp.sampler, rp.seed, rp.spp_x, rp.spp_y, pixel, k, g, 3u * d + 2u);
              float aux = pt_rng(g, rslot(d, SLOT_AUX), rp.seed);
              f3 wo = neg3(norm3(dir));
              Scatter sc;
              // guided: the learned mixture of this point (c = 0: not guided); the
              // technique (guide with probability c, else the BSDF) is drawn first
              const float *mix = nullptr;
              float gc = 0.f;
              bool use_guide = false;
              if constexpr (EST == PT_EST_GUIDED) {
                if (d < w.gverts) {
                  if (d == 0u && w.gpix) {
                    // the query the camera vertices of this pixel share; whether
                    // to guide follows this path's own material
                    mix = w.gmix + (size_t)(pixel - (uint32_t)(rp.chunk_first / rp.spp_count)) * PT_G_MIX;
                    gc = g_guidable(mat->kind, mat->p, w.gkinds) ? mix[40] : 0.f;
                  } else {
                    mix = w.gmix + (size_t)i * PT_G_MIX;
                    gc = mix[40];
                  }
                }
                use_guide = gc > 0.f && pt_rng(g, rslot(d, SLOT_GUIDE), rp.seed) < gc;
              }
              bool scat;
              if (use_guide) {
                sc.wi = g_sample(mix, pt_rng(g, rslot(d, SLOT_GUIDE + 1u), rp.seed),
                                 pt_rng(g, rslot(d, SLOT_GUIDE + 2u), rp.seed),
                                 pt_rng(g, rslot(d, SLOT_GUIDE + 3u), rp.seed));
                sc.flags = 0u;
                sc.att = s3(0.f);
                sc.pdf = 0.f;
                scat = true;
              } else {
                scat = bsdf_scatter<MS>(*mat, wo, h.sp, uc, u0, u1, aux, sc);
              }
              if (!scat) {
                L = add3(L, mul3(beta, bsdf_emit(*mat, dir, h.sp)));
                done = true;
              } else if (sc.flags & BX_SPECULAR) {
                beta = mul3(beta, sc.att);
                wi = sc.wi;
                cont = d + 1u < D;
                fl = 1u;
              } else {
                f3 wn = norm3(sc.wi);
                float cs = fabsf(dot3(h.sp.n, wn));
                if (EST == 0) {
                  // NEE with light_samples jittered offsets (maxdepth_nee.hpp:90-232)
                  if (N > 0) {
                    float inv_ns = 1.f / (float)rp.light_samples;
                    for (uint32_t si = 0; si < rp.light_samples && si < 4u; ++si) {
                      uint32_t ii = si / rp.ly, jj = si % rp.ly;
                      float a0 = div_rn((float)ii, (float)rp.lx), a1 = div_rn((float)(ii + 1u), (float)rp.lx);
                      float b0 = div_rn((float)jj, (float)rp.ly), b1 = div_rn((float)(jj + 1u), (float)rp.ly);
                      float ox = a0 + (a1 - a0) * pt_rng(g, rslot(d, SLOT_NEE_OFF + 2u * si), rp.seed);
                      float oy = b0 + (b1 - b0) * pt_rng(g, rslot(d, SLOT_NEE_OFF + 2u * si + 1u), rp.seed);
                      int li = (int)(pt_rng(g, rslot(d, SLOT_NEE_LIGHT + si), rp.seed) * (float)N) % N;
                      const DevLight &Lg = s.lights[li];
                      f3 nl;
                      f3 lp = light_point(Lg, ox, oy, nl);
                      f3 c = s3(0.f), lwi = s3(0.f);
                      float tmax = 0.f;
                      bool valid = false;
                      if (Lg.type == LIGHT_ENV) {
                        lwi = norm3(sub3(lp, p));
                        float pdf = (1.f / (float)N) * (1.f / (4.f * PT_PI));
                        float fac = fmaxf(0.f, dot3(h.sp.n, lwi));
                        f3 f = bsdf_f<MS>(*mat, wo, lwi, h.sp);
                        c = divs3(scl3(mul3(f, env_lookup(s, lwi)), fac), pdf);
                        tmax = 3.402823466e+38f;
                        valid = true;
                      } else {
                        f3 ld = sub3(lp, p);
                        float len = len3(ld);
                        lwi = divs3(ld, len);
                        if (len >= 0.001f) {
                          f3 le = (dot3(nl, neg3(lwi)) > 0.f) ? Lg.le : s3(0.f);
                          float pdf = (Lg.area <= 1e-7f) ? 0.f : 1.f / (Lg.area * (float)N);
                          float noi = fmaxf(0.f, dot3(h.sp.n, lwi));
                          float nol = fmaxf(0.f, dot3(nl, neg3(lwi)));
                          float d2 = len * len;
                          float fac = (d2 < 1e-7f) ? 0.f : (nol * noi) / d2;
                          if (pdf > 1e-7f) {
                            f3 f = bsdf_f<MS>(*mat, wo, lwi, h.sp);
                            c = divs3(scl3(mul3(f, le), fac), pdf);
                            tmax = len - 0.001f;
                            valid = true;
                          }
                        }
                      }
                      c = scl3(mul3(beta, c), inv_ns);
                      // zero contributions are dropped; non-finite ones are traced so the film's
                      // NaN guard sees them
                      if (valid && ((c.x > 0.f || c.y > 0.f || c.z > 0.f) || !finite3(c)))
                        emit_shadow(w, p, lwi, tmax, c, slot);
                    }
                  }
                } else if (N > 0) {
                  int li = (int)(pt_rng(g, rslot(d, SLOT_MIS_LIGHT), rp.seed) * (float)N) % N;
                  const DevLight &Lg = s.lights[li];
                  float e0 = pt_rng(g, rslot(d, SLOT_NEE_OFF), rp.seed);
                  float e1 = pt_rng(g, rslot(d, SLOT_NEE_OFF + 1u), rp.seed);
                  f3 lwi = s3(0.f), le = s3(0.f);
                  float pl = 0.f, tmax = 0.f;
                  if (Lg.type == LIGHT_ENV) {
                    float2 euv;
                    lwi = env_sample(s, e0, e1, pl, &euv);
                    le = euv.x >= 0.f ? env_lookup_uv(s, euv.x, euv.y) : env_lookup(s, lwi);
                    tmax = 3.402823466e+38f;
                  } else {
                    f3 nl;
                    f3 lp = light_point(Lg, e0, e1, nl);
                    f3 ld = sub3(lp, p);
                    float len = len3(ld);
                    lwi = divs3(ld, len);
                    float nol = dot3(nl, neg3(lwi));
                    if (len >= 0.001f && nol > 0.f && Lg.area > 0.f) {
                      pl = len * len / (Lg.area * nol);
                      le = Lg.le;
                      tmax = len - 0.001f;
                    }
                  }
                  // a reflection-only BSDF takes no light from behind its surface (the NEE
                  // integrator's max(0, n.wi), maxdepth_nee.hpp:137, 192)
                  if (bsdf_reflect_only(mat->kind) && dot3(h.sp.n, lwi) <= 0.f) pl = 0.f;
                  if (pl > 0.f) {
                    pl /= (float)N;
                    f3 f;
                    float pb;
                    bsdf_f_pdf<MS>(*mat, wo, lwi, h.sp, f, pb);
                    if (EST == PT_EST_GUIDED && gc > 0.f) pb = gc * g_pdf(mix, lwi) + (1.f - gc) * pb;
                    float cl = fabsf(dot3(h.sp.n, lwi));
                    f3 c = scl3(mul3(beta, mul3(f, le)), cl / pl * power_w(pl, pb));
                    if ((c.x > 0.f || c.y > 0.f || c.z > 0.f) || !finite3(c)) emit_shadow(w, p, lwi, tmax, c, slot);
                  }
                }
                float pm = sc.pdf;
                if (EST == PT_EST_GUIDED && gc > 0.f) {
                  // one-sample MIS over the two techniques (a guided direction behind a
                  // reflection-only surface carries nothing)
                  const bool below = bsdf_reflect_only(mat->kind) && dot3(h.sp.n, wn) <= 0.f;
                  f3 fg;
                  float pbs;
                  bsdf_f_pdf<MS>(*mat, wo, wn, h.sp, fg, pbs);
                  const f3 f = below ? s3(0.f) : fg;
                  pm = gc * g_pdf(mix, wn) + (1.f - gc) * pbs;
                  if (w.grec && d < w.grec_verts && slot % w.grec_stride == 0u) {
                    float *rec = w.grec + ((size_t)(slot / w.grec_stride) * w.grec_verts + d) * PT_G_REC;
                    // the network's inputs at this point (k_guide_infer's)
                    rec[0] = fminf(1.f, fmaxf(0.f, (p.x - w.glo[0]) * w.ginv[0]));
                    rec[1] = fminf(1.f, fmaxf(0.f, (p.y - w.glo[1]) * w.ginv[1]));
                    rec[2] = fminf(1.f, fmaxf(0.f, (p.z - w.glo[2]) * w.ginv[2]));
                    rec[3] = fminf(1.f, fmaxf(0.f, (h.sp.n.x + 1.f) * 0.5f));
                    rec[4] = fminf(1.f, fmaxf(0.f, (h.sp.n.y + 1.f) * 0.5f));
                    rec[5] = fminf(1.f, fmaxf(0.f, (h.sp.n.z + 1.f) * 0.5f));
                    rec[6] = wn.x; rec[7] = wn.y; rec[8] = wn.z;
                    rec[9] = pbs;
                    rec[10] = beta.x; rec[11] = beta.y; rec[12] = beta.z;
                    rec[16] = __uint_as_float(w.gpass);
                  }
                  beta = pm > 0.f ? scl3(mul3(beta, f), cs / pm) : s3(0.f);
                } else {
                  beta = scl3(mul3(beta, sc.att), cs / sc.pdf);
                }
                wi = wn;
                fl = 0u;
                if (EST == 0) {
                  cont = d + 1u < D;  // the emitter check of the next hit happens at the next vertex
                } else {
                  cont = d + 1u <= D;
                  if (EST == PT_EST_GUIDED) cont = cont && (beta.x != 0.f || beta.y != 0.f || beta.z != 0.f);
                  ppdf = pm;
                  pprev = p;
                }
              }
              if (cont && rp.rr_depth && d >= rp.rr_depth) {
                float qs = fminf(1.f, maxc3(beta));
                if (!(pt_rng(g, rslot(d, SLOT_RR), rp.seed) < qs)) cont = false;
                else beta = divs3(beta, qs);
              }
            }
          }
          // this path's slot is written by no other lane; this vertex's shadow rays
          // add to it only after this kernel
          if (L.x != 0.f || L.y != 0.f || L.z != 0.f || !finite3(L)) {
            w.lx[slot] += L.x;
            w.ly[slot] += L.y;
            w.lz[slot] += L.z;
          }
        }
        // converged: append the continuing paths to the same partition of the other buffer
        uint32_t j = base + wave_append(cout, cont);
        if (cont) {
          out.ox[j] = p.x; out.oy[j] = p.y; out.oz[j] = p.z;
          out.dx[j] = wi.x; out.dy[j] = wi.y; out.dz[j] = wi.z;
          out.bx[j] = beta.x; out.by[j] = beta.y; out.bz[j] = beta.z;
          out.flags[j] = fl;
          out.ppdf[j] = ppdf;
          out.px[j] = pprev.x; out.py[j] = pprev.y; out.pz[j] = pprev.z;
          out.slot[j] = slot;
          if (EST == PT_EST_RESERVOIR) {
            out.pbx[j] = pbeta.x; out.pby[j] = pbeta.y; out.pbz[j] = pbeta.z;
          }
        }
        // the list's first 256 entries are shaded: move the rest to the front
        const uint32_t rem = cnt > 256u ? cnt - 256u : 0u;
        __syncthreads();
        const uint32_t mv = threadIdx.x < rem ? o_list[c][256u + threadIdx.x] : 0u;
        __syncthreads();
        if (threadIdx.x < rem) o_list[c][threadIdx.x] = mv;
        if (threadIdx.x == 0) o_n[c] = rem;
        __syncthreads();
        cnt = rem;
      }
    }
    if (last) break;
    t0 = t_next;
  }
}

// ------------------------------------------------------------------ film
// Per-pixel sums of a batch chunk, PT_FILM_BATCH_T lanes per pixel (the
// renders of a chunk differ in samples per pixel, so T is fixed): lane j sums
// the pixel's samples j, j + T, ... and an xor-shuffle tree adds the T
// partial sums, a fixed order.
#define PT_FILM_BATCH_T 8u
__global__ void __launch_bounds__(256) k_film_batch(Wave w, unsigned long long *stats, BatchView b) {
  const uint32_t t = gtid(), j = t % PT_FILM_BATCH_T, pl = t / PT_FILM_BATCH_T;
  const ChunkInfo ci = b.chunks[*b.cur];
  const bool act = pl < ci.npix;  // the T lanes of a pixel agree
  float sr = 0.f, sg = 0.f, sb = 0.f;
  uint32_t nan = 0;
  uint64_t pixel = 0;
  float *film = nullptr;
  if (act) {
    const ChunkSeg sg0 = b.segs[find_seg_pix(b, ci, pl)];
    const uint32_t sppc = b.params[sg0.inst].spp_count;
    pixel = sg0.pix_first + (pl - sg0.pix_offset);
    const uint32_t slot0 = sg0.path_first + (pl - sg0.pix_offset) * sppc;
    film = b.films[sg0.inst];
    for (uint32_t k = j; k < sppc; k += PT_FILM_BATCH_T) {
      const uint32_t slot = slot0 + k;
      float r = w.lx[slot], g = w.ly[slot], bb = w.lz[slot];
      if (!(isfinite(r) && isfinite(g) && isfinite(bb))) {
        r = g = bb = 0.f;
        ++nan;
      }
      sr += r; sg += g; sb += bb;
    }
  }
  for (uint32_t off = PT_FILM_BATCH_T >> 1; off >= 1u; off >>= 1) {  // every lane of the wave takes part
    sr += __shfl_xor(sr, (int)off);
    sg += __shfl_xor(sg, (int)off);
    sb += __shfl_xor(sb, (int)off);
  }
  if (act && j == 0u) {
    film[3 * pixel] += sr;
    film[3 * pixel + 1] += sg;
    film[3 * pixel + 2] += sb;
  }
  if (nan) atomicAdd(&stats[CNT_NAN], (unsigned long long)nan);
}

// Per-pixel sums of a single render with T lanes per pixel (T a power of two,
// at most 64 and at most the pixel's samples): lane j sums the pixel's
// samples j, j + T, j + 2T, ... (so consecutive lanes read consecutive
// slots), then the T partial sums are added pairwise by an xor-shuffle tree.
// The order is fixed by (samples per pixel, T), so the film is as
// deterministic as the slots' radiance: with one shadow ray per vertex
// (NEE+MIS, guided) each slot takes one add per shadow pass and is
// reproducible bit for bit; the reference estimator with light_samples > 1 and
// the volume estimator add up to four unoccluded contributions per slot in
// k_shadow with float atomics in arbitrary order, so their slots (and films)
// can differ in the last place between runs.
// One lane per pixel strided through its pixel's samples and fetched 2.2 GB
// per 8M-slot chunk for 100 MB of radiance.
__global__ void __launch_bounds__(256) k_film_lanes(RenderParams rp, Wave w, float *film, float *samples,
                                                    unsigned long long *stats, uint32_t log2_t) {
  const uint32_t t = gtid(), T = 1u << log2_t;
  const uint32_t sppc = rp.spp_count, npix = rp.chunk_paths / sppc;
  const uint32_t pl = t >> log2_t, j = t & (T - 1u);
  const bool act = pl < npix;  // the T lanes of a pixel agree
  const uint64_t pixel = rp.chunk_first / sppc + pl;
  float sr = 0.f, sg = 0.f, sb = 0.f;
  uint32_t nan = 0;
  if (act) {
    const uint32_t slot0 = pl * sppc;
    for (uint32_t k = j; k < sppc; k += T) {
      const uint32_t slot = slot0 + k;
      float r = w.lx[slot], g = w.ly[slot], bb = w.lz[slot];
      if (!(isfinite(r) && isfinite(g) && isfinite(bb))) {
        r = g = bb = 0.f;
        ++nan;
      }
      if (samples) {
        const size_t o = 3 * ((size_t)pixel * sppc + k);
        samples[o] = r; samples[o + 1] = g; samples[o + 2] = bb;
      }
      sr += r; sg += g; sb += bb;
    }
  }
  for (uint32_t off = T >> 1; off >= 1u; off >>= 1) {  // every lane of the wave takes part
    sr += __shfl_xor(sr, (int)off);
    sg += __shfl_xor(sg, (int)off);
    sb += __shfl_xor(sb, (int)off);
  }
  if (act && j == 0u) {
    film[3 * pixel] += sr;
    film[3 * pixel + 1] += sg;
    film[3 * pixel + 2] += sb;
  }
  if (nan) atomicAdd(&stats[CNT_NAN], (unsigned long long)nan);
}

// Per-frame jittered pixel offsets (FrameBuffer::setOffsets,
// framebuffer.cpp:131-145), computed as the host computes them for pt_render:
// sample k = i * spp_y + j lies in stratum (i, j).
__global__ void k_pixel_offsets(uint32_t spp_x, uint32_t spp_y, uint32_t seed, float *off) {
  const uint32_t k = gtid(), spp = spp_x * spp_y;
  if (k >= spp) return;
  const uint32_t i = k / spp_y, j = k % spp_y;
  const float a0 = (float)i / (float)spp_x, a1 = (float)(i + 1) / (float)spp_x;
  const float b0 = (float)j / (float)spp_y, b1 = (float)(j + 1) / (float)spp_y;
  off[k] = (float)(a0 + (a1 - a0) * (double)pt_rng(0xffffffffu - k, 0u, seed ^ 0x5bd1e995u));
  off[spp + k] = (float)(b0 + (b1 - b0) * (double)pt_rng(0xffffffffu - k, 1u, seed ^ 0x5bd1e995u));
}

__global__ void k_film_scale(float *film, size_t n, float scale) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) film[i] *= scale;
}

// Advances the batch to its next chunk (the last launch of a chunk).
__global__ void k_next_chunk(BatchView b) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *b.cur += 1u;
}
__global__ void k_set_chunk(BatchView b, uint32_t c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *b.cur = c;
}

// ------------------------------------------------------------------ parity kernels
__global__ void k_sample(DevScene s, uint32_t sampler, uint32_t seed, uint32_t sx, uint32_t sy,
                         const uint32_t *dim, const uint64_t *index, const uint32_t *pixel, size_t n, float *out) {
  size_t i = gtid();
  if (i >= n) return;
  uint32_t px = pixel ? pixel[i] : 0u;
  out[i] = sample_dim(s, sampler, seed, sx, sy, px, (uint32_t)index[i], (uint32_t)index[i], dim[i]);
}

__global__ void __launch_bounds__(PT_BLOCK) k_intersect(DevScene s, const float *rays, size_t n, int any,
                                                        int32_t *prim, float *tout) {
  __shared__ int lds[PT_TRAV_LDS(PT_K_EXT, true) > PT_TRAV_LDS(PT_K_SHADOW, false)
                         ? PT_TRAV_LDS(PT_K_EXT, true)
                         : PT_TRAV_LDS(PT_K_SHADOW, false)];
  size_t i = gtid();
  if (i >= n) return;
  const float *r = rays + 8 * i;
  float t;
  int pr;
  uint32_t a = 0, b = 0;
  bool hit;
  if (any && s.qany)
    hit = traverse<true, true, false, PT_K_SHADOW>(s, lds, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), r[6], r[7], t, pr, a, b);
  else if (any)
    hit = traverse<true, false, false, PT_K_SHADOW>(s, lds, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), r[6], r[7], t, pr, a, b);
  else
    hit = traverse<false, false, false, PT_K_EXT>(s, lds, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), r[6], r[7], t, pr, a, b);
  prim[i] = hit ? (any ? 1 : pr) : -1;
  tout[i] = hit ? (any ? 0.f : t) : __int_as_float(0x7f800000);
}

__global__ void k_bsdf_eval(pt_material_desc m, const float *wo, const float *wi, const float *nrm,
                            const uint8_t *inside, size_t n, float *f, float *pdf) {
  size_t i = gtid();
  if (i >= n) return;
  ShadePoint h;
  h.n = mk3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  h.inside = inside ? inside[i] != 0 : true;
  f3 o = mk3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
  f3 w = mk3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
  f3 r;
  float p;
  bsdf_f_pdf(m, o, w, h, r, p);  // the renderer's evaluation of a light direction
  f[3 * i] = r.x; f[3 * i + 1] = r.y; f[3 * i + 2] = r.z;
  pdf[i] = p;
}

__global__ void k_bsdf_sample(pt_material_desc m, const float *wo, const float *nrm, const uint8_t *inside,
                              const float *uc, const float *u2, const float *aux, size_t n, float *wi, float *f,
                              float *pdf, uint32_t *flags, uint8_t *ok) {
  size_t i = gtid();
  if (i >= n) return;
  ShadePoint h;
  h.n = mk3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  h.inside = inside ? inside[i] != 0 : true;
  f3 o = mk3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
  Scatter sc;
  bool r = bsdf_scatter(m, norm3(o), h, uc[i], u2[2 * i], u2[2 * i + 1], aux ? aux[i] : 0.5f, sc);
  ok[i] = r ? 1 : 0;
  wi[3 * i] = sc.wi.x; wi[3 * i + 1] = sc.wi.y; wi[3 * i + 2] = sc.wi.z;
  f[3 * i] = sc.att.x; f[3 * i + 1] = sc.att.y; f[3 * i + 2] = sc.att.z;
  pdf[i] = sc.pdf;
  flags[i] = sc.flags;
}

__global__ void k_light_sample(DevScene s, uint32_t light, const float *u2, const float *ref, size_t n, float *p,
                               float *nrm, float *pdf, float *le) {
  size_t i = gtid();
  if (i >= n) return;
  const DevLight &L = s.lights[light];
  f3 nl;
  f3 pt = light_point(L, u2[2 * i], u2[2 * i + 1], nl);
  f3 o = mk3(ref[3 * i], ref[3 * i + 1], ref[3 * i + 2]);
  f3 w = norm3(sub3(pt, o));
  float pd;
  f3 e;
  if (L.type == LIGHT_ENV) {
    pd = 1.f / (4.f * PT_PI);
    e = env_lookup(s, w);
  } else if (L.type == LIGHT_QUAD) {
    f3 v = sub3(pt, o);
    pd = dot3(v, v) / (L.area * fabsf(dot3(L.n, neg3(w))));
    e = (dot3(nl, neg3(w)) > 0.f) ? L.le : s3(0.f);
  } else {
    f3 wc = sub3(L.c, o);
    float dc2 = dot3(wc, wc), r2 = L.r * L.r;
    pd = (dc2 <= r2) ? 1.f / (4.f * PT_PI) : 1.f / (2.f * PT_PI * (1.f - sqrtf(1.f - r2 / dc2)));
    e = (dot3(nl, neg3(w)) > 0.f) ? L.le : s3(0.f);
  }
  p[3 * i] = pt.x; p[3 * i + 1] = pt.y; p[3 * i + 2] = pt.z;
  nrm[3 * i] = nl.x; nrm[3 * i + 1] = nl.y; nrm[3 * i + 2] = nl.z;
  pdf[i] = pd;
  le[3 * i] = e.x; le[3 * i + 1] = e.y; le[3 * i + 2] = e.z;
}

__global__ void k_env_sample(DevScene s, const float *u2, size_t n, float *dir, float *pdf) {
  size_t i = gtid();
  if (i >= n) return;
  float pd;
  const f3 d = env_sample(s, u2[2 * i], u2[2 * i + 1], pd);
  dir[3 * i] = d.x; dir[3 * i + 1] = d.y; dir[3 * i + 2] = d.z;
  pdf[i] = pd;
}

// ------------------------------------------------------------------ launchers
static inline uint32_t blocks_for(size_t n, uint32_t bs = 256) { return (uint32_t)((n + bs - 1) / bs); }
// Grid of a partitioned kernel: enough blocks for max_n entries spread over the
// partitions, at most grid_cap, and a multiple of PT_PARTS so that every
// partition gets the same number of blocks.
static inline uint32_t part_grid(size_t max_n, uint32_t grid_cap) {
  uint32_t per_part = blocks_for((max_n + PT_PARTS - 1) / PT_PARTS);
  uint32_t g = std::min<uint64_t>((uint64_t)per_part * PT_PARTS, grid_cap);
  return std::max<uint32_t>(PT_PARTS, g / PT_PARTS * PT_PARTS);
}

// The render launchers take a BatchView: null renders the one scene of the
// arguments, non-null renders the batch's current chunk (max_n / the chunk
// fields of rp then only size the grids).
void launch_generate(hipStream_t st, const DevScene &s, const DevCamera &c, const RenderParams &rp, const Wave &w,
                     const float *offx, const float *offy, const BatchView *b) {
  const size_t per = (rp.chunk_paths + PT_PART_RUN * PT_PARTS - 1) / (PT_PART_RUN * PT_PARTS) * PT_PART_RUN;
  const size_t threads = std::max<size_t>(per * PT_PARTS, rp.chunk_paths);
  if (b)
    hipLaunchKernelGGL(k_generate<true>, dim3(blocks_for(threads)), dim3(256), 0, st, s, c, rp, w, offx, offy, *b);
  else
    hipLaunchKernelGGL(k_generate<false>, dim3(blocks_for(threads)), dim3(256), 0, st, s, c, rp, w, offx, offy,
                       BatchView{});
}
void launch_extend(hipStream_t st, const DevScene &s, const Wave &w, int qi, uint32_t max_n, uint32_t grid_cap,
                   bool count, const BatchView *b) {
  uint32_t g = part_grid(max_n, grid_cap);
  const BatchView bv = b ? *b : BatchView{};
  if (b) {
    if (count)
      hipLaunchKernelGGL((k_extend<true, true>), dim3(g), dim3(PT_BLOCK), 0, st, s, w, qi, bv);
    else
      hipLaunchKernelGGL((k_extend<false, true>), dim3(g), dim3(PT_BLOCK), 0, st, s, w, qi, bv);
  } else {
    if (count)
      hipLaunchKernelGGL((k_extend<true, false>), dim3(g), dim3(PT_BLOCK), 0, st, s, w, qi, bv);
    else
      hipLaunchKernelGGL((k_extend<false, false>), dim3(g), dim3(PT_BLOCK), 0, st, s, w, qi, bv);
  }
}
template <int EST, bool BATCH>
static void launch_shade_est(hipStream_t st, const DevScene &s, const RenderParams &rp, const Wave &w, int qi,
                             uint32_t d, uint32_t g, const BatchView &b, uint32_t mat_kinds) {
  // the smallest material set that holds every kind the scene uses (fewer
  // registers, less code); the microfacet set for the two estimators the
  // workloads render with
  constexpr bool micro = EST == PT_EST_NEE_MIS || EST == PT_EST_REFERENCE_NEE;
  if (micro && (mat_kinds & ~kMatMicro) == 0u)
    hipLaunchKernelGGL((k_shade<EST, micro ? kMatMicro : kMatBasic, BATCH>), dim3(g), dim3(256), 0, st, s, rp, w, qi,
                       d, b);
  else if ((mat_kinds & ~kMatBasic) == 0u)
    hipLaunchKernelGGL((k_shade<EST, kMatBasic, BATCH>), dim3(g), dim3(256), 0, st, s, rp, w, qi, d, b);
  else
    hipLaunchKernelGGL((k_shade<EST, kMatAll, BATCH>), dim3(g), dim3(256), 0, st, s, rp, w, qi, d, b);
}
template <bool BATCH>
static void launch_shade_b(hipStream_t st, const DevScene &s, const RenderParams &rp, const Wave &w, int qi,
                           uint32_t d, uint32_t g, const BatchView &b, uint32_t mat_kinds) {
  switch (rp.estimator) {
    case PT_EST_NEE_MIS: launch_shade_est<PT_EST_NEE_MIS, BATCH>(st, s, rp, w, qi, d, g, b, mat_kinds); break;
    case PT_EST_NAIVE: launch_shade_est<PT_EST_NAIVE, BATCH>(st, s, rp, w, qi, d, g, b, mat_kinds); break;
    case PT_EST_ROULETTE: launch_shade_est<PT_EST_ROULETTE, BATCH>(st, s, rp, w, qi, d, g, b, mat_kinds); break;
    case PT_EST_RESERVOIR: launch_shade_est<PT_EST_RESERVOIR, BATCH>(st, s, rp, w, qi, d, g, b, mat_kinds); break;
    case PT_EST_VOLUME: launch_shade_est<PT_EST_VOLUME, BATCH>(st, s, rp, w, qi, d, g, b, mat_kinds); break;
    case PT_EST_GUIDED:
      if constexpr (!BATCH) launch_shade_est<PT_EST_GUIDED, false>(st, s, rp, w, qi, d, g, b, mat_kinds);
      break;
    default: launch_shade_est<PT_EST_REFERENCE_NEE, BATCH>(st, s, rp, w, qi, d, g, b, mat_kinds); break;
  }
}
void launch_shade(hipStream_t st, const DevScene &s, const RenderParams &rp, const Wave &w, int qi, uint32_t d,
                  uint32_t max_n, uint32_t grid_cap, const BatchView *b, uint32_t mat_kinds) {
  uint32_t g = part_grid(max_n, grid_cap);
  if (b)
    launch_shade_b<true>(st, s, rp, w, qi, d, g, *b, mat_kinds);
  else
    launch_shade_b<false>(st, s, rp, w, qi, d, g, BatchView{}, mat_kinds);
}
void launch_shadow(hipStream_t st, const DevScene &s, const Wave &w, uint32_t max_n, uint32_t grid_cap, bool count,
                   bool visibility, const BatchView *b) {
  uint32_t g = part_grid(max_n, grid_cap);
  const BatchView bv = b ? *b : BatchView{};
  // any hit reads the quantised nodes when the scene (batched: every scene of
  // the batch) selects them
  const bool qn = s.qany != 0u;
#define PT_SHADOW_LAUNCH(C, V, B)                                                                     \
  do {                                                                                                \
    if (qn)                                                                                           \
      hipLaunchKernelGGL((k_shadow<C, V, B, true>), dim3(g), dim3(PT_BLOCK), 0, st, s, w, bv);       \
    else                                                                                              \
      hipLaunchKernelGGL((k_shadow<C, V, B, false>), dim3(g), dim3(PT_BLOCK), 0, st, s, w, bv);      \
  } while (0)
  if (b) {
    if (visibility) {
      if (count) PT_SHADOW_LAUNCH(true, true, true); else PT_SHADOW_LAUNCH(false, true, true);
    } else {
      if (count) PT_SHADOW_LAUNCH(true, false, true); else PT_SHADOW_LAUNCH(false, false, true);
    }
  } else {
    if (visibility) {
      if (count) PT_SHADOW_LAUNCH(true, true, false); else PT_SHADOW_LAUNCH(false, true, false);
    } else {
      if (count) PT_SHADOW_LAUNCH(true, false, false); else PT_SHADOW_LAUNCH(false, false, false);
    }
  }
#undef PT_SHADOW_LAUNCH
}
void launch_film(hipStream_t st, const RenderParams &rp, const Wave &w, float *film, float *samples,
                 unsigned long long *stats, const BatchView *b, uint32_t max_pix) {
  if (b) {
    hipLaunchKernelGGL(k_film_batch, dim3(blocks_for((size_t)max_pix * PT_FILM_BATCH_T)), dim3(256), 0, st, w, stats,
                       *b);
  } else {
    const uint32_t npix = rp.chunk_paths / rp.spp_count;
    uint32_t lg = 0;
    while (lg < 6u && (2u << lg) <= rp.spp_count) ++lg;  // T = 2^lg <= min(spp, 64)
    hipLaunchKernelGGL(k_film_lanes, dim3(blocks_for((size_t)npix << lg)), dim3(256), 0, st, rp, w, film, samples,
                       stats, lg);
  }
}
void launch_pixel_offsets(hipStream_t st, uint32_t spp_x, uint32_t spp_y, uint32_t seed, float *off) {
  hipLaunchKernelGGL(k_pixel_offsets, dim3(blocks_for((size_t)spp_x * spp_y)), dim3(256), 0, st, spp_x, spp_y, seed,
                     off);
}
void launch_film_scale(hipStream_t st, float *film, size_t n, float scale) {
  hipLaunchKernelGGL(k_film_scale, dim3(blocks_for(n)), dim3(256), 0, st, film, n, scale);
}
void launch_next_chunk(hipStream_t st, const BatchView &b) {
  hipLaunchKernelGGL(k_next_chunk, dim3(1), dim3(64), 0, st, b);
}
void launch_set_chunk(hipStream_t st, const BatchView &b, uint32_t c) {
  hipLaunchKernelGGL(k_set_chunk, dim3(1), dim3(64), 0, st, b, c);
}
void launch_sample(hipStream_t st, const DevScene &s, uint32_t sampler, uint32_t seed, uint32_t sx, uint32_t sy,
                   const uint32_t *dim, const uint64_t *index, const uint32_t *pixel, size_t n, float *out) {
  hipLaunchKernelGGL(k_sample, dim3(blocks_for(n)), dim3(256), 0, st, s, sampler, seed, sx, sy, dim, index, pixel,
                     n, out);
}
void launch_intersect(hipStream_t st, const DevScene &s, const float *rays, size_t n, int any, int32_t *prim,
                      float *t) {
  hipLaunchKernelGGL(k_intersect, dim3(blocks_for(n)), dim3(PT_BLOCK), 0, st, s, rays, n, any, prim, t);
}
void launch_bsdf_eval(hipStream_t st, const pt_material_desc &m, const float *wo, const float *wi, const float *nrm,
                      const uint8_t *inside, size_t n, float *f, float *pdf) {
  hipLaunchKernelGGL(k_bsdf_eval, dim3(blocks_for(n)), dim3(256), 0, st, m, wo, wi, nrm, inside, n, f, pdf);
}
void launch_bsdf_sample(hipStream_t st, const pt_material_desc &m, const float *wo, const float *nrm,
                        const uint8_t *inside, const float *uc, const float *u2, const float *aux, size_t n, float *wi,
                        float *f, float *pdf, uint32_t *flags, uint8_t *ok) {
  hipLaunchKernelGGL(k_bsdf_sample, dim3(blocks_for(n)), dim3(256), 0, st, m, wo, nrm, inside, uc, u2, aux, n, wi, f,
                     pdf, flags, ok);
}
void launch_light_sample(hipStream_t st, const DevScene &s, uint32_t light, const float *u2, const float *ref,
                         size_t n, float *p, float *nrm, float *pdf, float *le) {
  hipLaunchKernelGGL(k_light_sample, dim3(blocks_for(n)), dim3(256), 0, st, s, light, u2, ref, n, p, nrm, pdf, le);
}
void launch_env_sample(hipStream_t st, const DevScene &s, const float *u2, size_t n, float *dir, float *pdf) {
  hipLaunchKernelGGL(k_env_sample, dim3(blocks_for(n)), dim3(256), 0, st, s, u2, n, dir, pdf);
}

}  // namespace pt
