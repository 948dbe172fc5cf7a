// pt_guide.hip -- kernels of the learned path-guiding proposal (network,
// mixture and training; see pt_guide.h for the model and the layouts).
//
//   k_guide_infer    per live path of a vertex: shading point -> encoding ->
//                    MLP (MFMA, 64 points per wave) -> processed mixture
//   k_guide_points   the same network on caller-given inputs (raw outputs)
//   k_guide_mix      mixture pdf / samples of caller-given inputs
//   k_guide_train_a  per training sample: forward with stored activations,
//                    loss gradient at the outputs, backward through the
//                    layers (MFMA, 16 samples per wave); writes the layer
//                    inputs and output gradients, transposed, for ...
//   k_guide_train_b  weight gradients dW_l = X_l^T dZ_l over the batch
//                    (MFMA, split over the batch, partial sums per split)
//   k_guide_update   Adam on the FP32 master weights, then the BF16
//                    fragments (or packing alone)
//   k_guide_collect  training samples from the renderer's path records
#include "pt_device.h"
#include "pt_guide.h"
#include "pt_guide_launch.h"

namespace pt {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ uint32_t g_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Two C tiles (rows 16 (2s) + 4g + j and 16 (2s+1) + 4g + j) -> the B operand
// of k-step s of the next product; RELU applies the activation.
template <bool RELU>
PT_D bf16x8 to_operand(f32x4 lo, f32x4 hi) {
  bf16x8 r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    r[j] = (__bf16)(RELU ? fmaxf(lo[j], 0.f) : lo[j]);
    r[4 + j] = (__bf16)(RELU ? fmaxf(hi[j], 0.f) : hi[j]);
  }
  return r;
}

// One-blob features of input value x for lane group g: bins 4g..4g+3 and
// 16+4g..16+4g+3 (the k-step order of g_kperm).
PT_D bf16x8 encode(float x, uint32_t g) {
  const float y = 32.f * x - 0.5f;
  bf16x8 r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    r[j] = (__bf16)g_bump((float)(4u * g + j) - y);
    r[4 + j] = (__bf16)g_bump((float)(16u + 4u * g + j) - y);
  }
  return r;
}

PT_D f32x4 bias_tile(const float *bias, int l, uint32_t nt, uint32_t g) {
  f32x4 b;
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = bias[l * 64 + 16 * nt + 4 * g + j];
  return b;
}

// The network on T tiles of 16 points: xin[t][v] are lane (q, g)'s point q of
// tile t; wf are the forward fragments (LDS), bias the padded biases.
// Returns the output tiles acc[t][nt] (rows 16 nt + 4 g + j of point q).
template <int T>
PT_D void g_network(const bf16x8 *wf, const float *bias, const float (&xin)[T][PT_G_IN], uint32_t lane,
                    f32x4 (&acc)[T][4]) {
  const uint32_t g = lane >> 4;
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt) acc[t][nt] = bias_tile(bias, 0, nt, g);
#pragma unroll
  for (uint32_t s = 0; s < 6; ++s) {
    bf16x8 bop[T];
#pragma unroll
    for (int t = 0; t < T; ++t) bop[t] = encode(xin[t][s], g);
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt) {
      const bf16x8 a = wf[(nt * 6 + s) * 64 + lane];
#pragma unroll
      for (int t = 0; t < T; ++t) acc[t][nt] = mfma(a, bop[t], acc[t][nt]);
    }
  }
#pragma unroll
  for (int l = 1; l < PT_G_LAYERS; ++l) {
    bf16x8 in[T][2];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      in[t][0] = to_operand<true>(acc[t][0], acc[t][1]);
      in[t][1] = to_operand<true>(acc[t][2], acc[t][3]);
    }
    const bf16x8 *wl = wf + g_fwd_frag_offset(l) / 8;
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
      for (uint32_t nt = 0; nt < 4; ++nt) acc[t][nt] = bias_tile(bias, l, nt, g);
#pragma unroll
    for (uint32_t s = 0; s < 2; ++s)
#pragma unroll
      for (uint32_t nt = 0; nt < 4; ++nt) {
        const bf16x8 a = wl[(nt * 2 + s) * 64 + lane];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t][nt] = mfma(a, in[t][s], acc[t][nt]);
      }
  }
}

// Mixture parameters of one point from its output tiles (lane (q, g)): lanes
// of groups 0 and 1 get lobes 4g..4g+3, group 2 the selection probability.
struct LobeQuad {
  float kappa[4], mux[4], muy[4], muz[4], pi[4], b[4], v[4];
  float inv_len[4], cos_mw[4];
  float c, s_sig;
};
PT_D void g_process(const f32x4 (&o)[4], uint32_t g, float c_lo, float c_span, LobeQuad &r) {
  // y components of lobes 4g + j sit in tile 2 of group g + 2
  float my[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) my[j] = __shfl_xor(o[2][j], 32);
  float amax = -INFINITY;
#pragma unroll
  for (int j = 0; j < 4; ++j) amax = fmaxf(amax, o[0][j]);
  amax = fmaxf(amax, __shfl_xor(amax, 16));
  float e[4], sum = 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    e[j] = expf(o[0][j] - amax);
    sum += e[j];
  }
  sum += __shfl_xor(sum, 16);
  const float s_sel = __shfl(o[0][0], (int)((threadIdx.x & 15u) | 32u));  // row 8: group 2, j = 0
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    r.kappa[j] = g_kappa(o[1][j]);
    float l = sqrtf(o[2][j] * o[2][j] + my[j] * my[j] + o[3][j] * o[3][j]);
    l = fmaxf(l, 1e-12f);
    r.inv_len[j] = 1.f / l;
    r.mux[j] = o[2][j] * r.inv_len[j];
    r.muy[j] = my[j] * r.inv_len[j];
    r.muz[j] = o[3][j] * r.inv_len[j];
    r.pi[j] = e[j] / sum;
    r.b[j] = logf(r.pi[j]) + g_log_norm(r.kappa[j]) - r.kappa[j];
  }
  const float sg = 1.f / (1.f + expf(-s_sel));
  r.s_sig = sg;
  r.c = c_lo + c_span * sg;
  (void)g;
}

// ------------------------------------------------------------------ inference
#define PT_GQ_STRIDE 7  // per point in LDS: 6 inputs + guided flag
// A block of PT_GI_BLOCK threads shares one LDS copy of the weights; each of
// its waves evaluates PT_GI_T tiles of 16 points, so a block iteration covers
// PT_GI_PTS points.  Two tiles per wave keep the kernel within 128 registers
// (4 waves per SIMD; four tiles needed 202), and 512-thread blocks keep the
// weights' 56 KB of LDS to two copies per CU.
#ifndef PT_GI_BLOCK
#define PT_GI_BLOCK 512
#endif
#ifndef PT_GI_T
#define PT_GI_T 2
#endif
#define PT_GI_PTS (PT_GI_BLOCK / 64 * 16 * PT_GI_T)
static_assert(PT_GI_PTS <= PT_GI_BLOCK, "one input thread per point");
// With pix (the camera vertex of a guided render), the items are the
// chunk's pixels rather than its live entries: pixel i of the chunk is
// evaluated at the hit of its first sample in the chunk, found in the queue
// where k_generate placed that slot, and its mixture goes to gmix[48 i]
// (a sample that hit an unguided material still gives the pixel a mixture,
// from its point's inputs; one that missed, from zero inputs).
struct InferPixels {
  uint64_t chunk_first;  // the chunk's first path sample
  uint32_t sppc, n;      // samples per pixel of the pass, path samples of the chunk
  uint32_t on;
};
PT_D uint32_t entry_of_slot(const Wave &w, uint32_t slot) {  // k_generate's slot -> entry layout
  const uint32_t run = slot / PT_PART_RUN;
  return (run % PT_PARTS) * w.part_cap + (run / PT_PARTS) * PT_PART_RUN + slot % PT_PART_RUN;
}
__global__ void __launch_bounds__(PT_GI_BLOCK) k_guide_infer(DevScene s, Wave w, int qi, GuideView gv, uint32_t kinds,
                                                             InferPixels px) {
  __shared__ bf16x8 wf[PT_G_FWD_VALUES / 8];
  __shared__ float bias[PT_G_BIAS_VALUES];
  __shared__ float qin[PT_GI_PTS * PT_GQ_STRIDE];
  const uint32_t tid = threadIdx.x;
  {
    const uint4 *src = reinterpret_cast<const uint4 *>(gv.fwd);
    uint4 *dst = reinterpret_cast<uint4 *>(wf);
    for (uint32_t i = tid; i < PT_G_FWD_VALUES / 8; i += PT_GI_BLOCK) dst[i] = src[i];
    for (uint32_t i = tid; i < PT_G_BIAS_VALUES; i += PT_GI_BLOCK) bias[i] = gv.bias[i];
  }
  __syncthreads();
  const uint32_t part = blockIdx.x % PT_PARTS, nb = gridDim.x / PT_PARTS;
  const uint64_t pix0 = px.on ? px.chunk_first / px.sppc : 0u;
  // items: the partition's live entries, or the chunk's pixels (all blocks)
  const uint32_t n = px.on ? (uint32_t)((px.chunk_first + px.n - 1u) / px.sppc - pix0 + 1u) : *w.live(qi, part);
  const uint32_t base = px.on ? 0u : part * w.part_cap;
  const uint32_t first = px.on ? blockIdx.x * PT_GI_PTS : (blockIdx.x / PT_PARTS) * PT_GI_PTS;
  const uint32_t step = px.on ? gridDim.x * PT_GI_PTS : nb * PT_GI_PTS;
  const PathState &ps = w.st[qi];
  const uint32_t lane = tid & 63u, wv = tid >> 6, q = lane & 15u, g = lane >> 4;
  for (uint32_t i0 = first; i0 < n; i0 += step) {
    if (tid < PT_GI_PTS) {
      const uint32_t i = i0 + tid;
      float x[PT_G_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      float ok = 0.f;
      if (i < n) {
        uint32_t e = base + i;
        if (px.on) {
          const uint64_t q0 = (pix0 + i) * px.sppc;
          e = entry_of_slot(w, (uint32_t)((q0 > px.chunk_first ? q0 : px.chunk_first) - px.chunk_first));
        }
        const int prim = ps.prim[e];
        if (prim >= 0) {
          const f3 o = mk3(ps.ox[e], ps.oy[e], ps.oz[e]);
          const f3 d = mk3(ps.dx[e], ps.dy[e], ps.dz[e]);
          const SurfHit h = surface(s, prim, o, d, ps.t[e]);
          const pt_material_desc &m = s.mats[h.mat];
          ok = g_guidable(m.kind, m.p, kinds) ? 1.f : 0.f;
          if (ok != 0.f || px.on) {
            x[0] = fminf(1.f, fmaxf(0.f, (h.p.x - gv.lo[0]) * gv.inv_ext[0]));
            x[1] = fminf(1.f, fmaxf(0.f, (h.p.y - gv.lo[1]) * gv.inv_ext[1]));
            x[2] = fminf(1.f, fmaxf(0.f, (h.p.z - gv.lo[2]) * gv.inv_ext[2]));
            x[3] = fminf(1.f, fmaxf(0.f, (h.sp.n.x + 1.f) * 0.5f));
            x[4] = fminf(1.f, fmaxf(0.f, (h.sp.n.y + 1.f) * 0.5f));
            x[5] = fminf(1.f, fmaxf(0.f, (h.sp.n.z + 1.f) * 0.5f));
          }
        }
      }
#pragma unroll
      for (int v = 0; v < PT_G_IN; ++v) qin[tid * PT_GQ_STRIDE + v] = x[v];
      qin[tid * PT_GQ_STRIDE + 6] = ok;
    }
    __syncthreads();
    float xin[PT_GI_T][PT_G_IN];
#pragma unroll
    for (int t = 0; t < PT_GI_T; ++t)
#pragma unroll
      for (int v = 0; v < PT_G_IN; ++v) xin[t][v] = qin[(16u * PT_GI_T * wv + 16u * t + q) * PT_GQ_STRIDE + v];
    f32x4 acc[PT_GI_T][4];
    g_network<PT_GI_T>(wf, bias, xin, lane, acc);
#pragma unroll
    for (int t = 0; t < PT_GI_T; ++t) {
      LobeQuad r;
      g_process(acc[t], g, gv.c_lo, gv.c_span, r);
      const uint32_t pl = 16u * PT_GI_T * wv + 16u * t + q;  // point of the block
      const uint32_t i = i0 + pl;
      if (i < n) {
        const bool ok = qin[pl * PT_GQ_STRIDE + 6] != 0.f;
        float4 *mix = reinterpret_cast<float4 *>(w.gmix + (size_t)(base + i) * PT_G_MIX);
        if (g < 2u && (ok || px.on)) {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            mix[4 * g + j] = make_float4(r.kappa[j] * r.mux[j], r.kappa[j] * r.muy[j], r.kappa[j] * r.muz[j], r.b[j]);
          mix[8 + g] = make_float4(r.pi[0], r.pi[1], r.pi[2], r.pi[3]);
        } else if (g == 2u) {
          // c; per entry, 0 where the point is not guided (a pixel's shared
          // query keeps c: each path decides by its own material)
          mix[10] = make_float4(ok || px.on ? r.c : 0.f, 0.f, 0.f, 0.f);
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ caller-given points
// raw (n x 41) outputs and/or processed mixtures (n x 48) of n points x (n x 6).
__global__ void __launch_bounds__(256) k_guide_points(GuideView gv, const float *xs, size_t n, float *raw,
                                                      float *mixout) {
  __shared__ bf16x8 wf[PT_G_FWD_VALUES / 8];
  __shared__ float bias[PT_G_BIAS_VALUES];
  const uint32_t tid = threadIdx.x;
  {
    const uint4 *src = reinterpret_cast<const uint4 *>(gv.fwd);
    uint4 *dst = reinterpret_cast<uint4 *>(wf);
    for (uint32_t i = tid; i < PT_G_FWD_VALUES / 8; i += 256) dst[i] = src[i];
    for (uint32_t i = tid; i < PT_G_BIAS_VALUES; i += 256) bias[i] = gv.bias[i];
  }
  __syncthreads();
  const uint32_t lane = tid & 63u, wv = tid >> 6, q = lane & 15u, g = lane >> 4;
  const size_t p0 = (size_t)blockIdx.x * 256u + 64u * wv;
  float xin[4][PT_G_IN];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const size_t p = p0 + 16u * t + q;
#pragma unroll
    for (int v = 0; v < PT_G_IN; ++v) xin[t][v] = p < n ? xs[p * PT_G_IN + v] : 0.f;
  }
  f32x4 acc[4][4];
  g_network<4>(wf, bias, xin, lane, acc);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const size_t p = p0 + 16u * t + q;
    if (raw && p < n) {
      for (uint32_t o = 0; o < PT_G_OUT; ++o) {
        const uint32_t row = g_out_row(o);
        if ((row & 15u) / 4u == g) raw[(size_t)p * PT_G_OUT + o] = acc[t][row / 16u][row & 3u];
      }
    }
    LobeQuad r;
    g_process(acc[t], g, gv.c_lo, gv.c_span, r);
    if (mixout && p < n) {
      float *mx = mixout + (size_t)p * PT_G_MIX;
      if (g < 2u) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t i = 4u * g + j;
          mx[4 * i] = r.kappa[j] * r.mux[j];
          mx[4 * i + 1] = r.kappa[j] * r.muy[j];
          mx[4 * i + 2] = r.kappa[j] * r.muz[j];
          mx[4 * i + 3] = r.b[j];
          mx[32 + i] = r.pi[j];
        }
      } else if (g == 2u) {
        mx[40] = r.c;
        for (int k = 41; k < PT_G_MIX; ++k) mx[k] = 0.f;
      }
    }
  }
}

// Guide pdf at given directions, and guide samples from given uniforms, of
// processed mixtures (the renderer's device functions).
__global__ void k_guide_mix(const float *mix, const float *dirs, const float *u, size_t n, float *pdf,
                            float *out_dir, float *out_pdf) {
  const size_t i = g_tid();
  if (i >= n) return;
  const float *m = mix + i * PT_G_MIX;
  if (dirs) pdf[i] = g_pdf(m, mk3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
  if (u) {
    f3 d = g_sample(m, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    out_dir[3 * i] = d.x;
    out_dir[3 * i + 1] = d.y;
    out_dir[3 * i + 2] = d.z;
    out_pdf[i] = g_pdf(m, d);
  }
}

// ------------------------------------------------------------------ training
// Sample record: x (6), direction (3), p_bsdf, weight F/q, pad (12 floats).
// Loss per sample: -w (log p_g + log(c p_g + (1 - c) p_bsdf)), averaged.
#ifndef PT_G_SAMPLE
#define PT_G_SAMPLE 12
#endif
PT_D uint32_t g_hash(uint32_t a, uint32_t b) {
  uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u);
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

// Transposed activation / gradient store: row r of layer l's block, column b.
PT_D void store_col(uint16_t *base, uint32_t ld, uint32_t r, uint32_t b, __bf16 v) {
  base[(size_t)r * ld + b] = __builtin_bit_cast(uint16_t, v);
}

// Sum over the 16 points of a tile (lanes with equal g) of a C tile, into
// the lanes with q == 0.
PT_D f32x4 tile_colsum(f32x4 v) {
#pragma unroll
  for (int off = 1; off < 16; off <<= 1)
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] += __shfl_xor(v[j], off);
  return v;
}


__global__ void __launch_bounds__(256) k_guide_train_a(GuideView gv, const float *samples, uint32_t n_samples,
                                                       uint32_t batch, uint32_t ld, uint32_t pick_seed,
                                                       uint16_t *xt, uint16_t *dzt, float *dbpart,
                                                       float *losspart) {
  __shared__ bf16x8 wf[PT_G_FWD_VALUES / 8];
  __shared__ bf16x8 wb[PT_G_BWD_VALUES / 8];
  __shared__ float bias[PT_G_BIAS_VALUES];
  const uint32_t tid = threadIdx.x;
  {
    const uint4 *src = reinterpret_cast<const uint4 *>(gv.fwd);
    uint4 *dst = reinterpret_cast<uint4 *>(wf);
    for (uint32_t i = tid; i < PT_G_FWD_VALUES / 8; i += 256) dst[i] = src[i];
    const uint4 *srcb = reinterpret_cast<const uint4 *>(gv.bwd);
    uint4 *dstb = reinterpret_cast<uint4 *>(wb);
    for (uint32_t i = tid; i < PT_G_BWD_VALUES / 8; i += 256) dstb[i] = srcb[i];
    for (uint32_t i = tid; i < PT_G_BIAS_VALUES; i += 256) bias[i] = gv.bias[i];
  }
  __syncthreads();
  const uint32_t lane = tid & 63u, q = lane & 15u, g = lane >> 4;
  const uint32_t wave = blockIdx.x * 4u + (tid >> 6);  // tile of 16 samples
  const uint32_t b = wave * 16u + q;                    // column of this lane's sample (< ld)
  // the lane's sample
  float x[PT_G_IN], wx = 0.f, wy = 0.f, wz = 1.f, pb = 0.f, wt = 0.f;
  for (int v = 0; v < PT_G_IN; ++v) x[v] = 0.f;
  if (b < batch) {
    const uint32_t idx = pick_seed ? g_hash(pick_seed, b) % n_samples : b;
    const float *sp = samples + (size_t)idx * PT_G_SAMPLE;
    for (int v = 0; v < PT_G_IN; ++v) x[v] = sp[v];
    wx = sp[6]; wy = sp[7]; wz = sp[8];
    pb = sp[9];
    wt = sp[10];
  }
  const float inv_batch = 1.f / (float)batch;
  // -- forward, storing the layer inputs and the ReLU masks
  f32x4 acc[4];
  uint32_t mask[4] = {0u, 0u, 0u, 0u};  // bit 4 nt + j: unit 16 nt + 4 g + j of layer l active
#pragma unroll
  for (uint32_t nt = 0; nt < 4; ++nt) acc[nt] = bias_tile(bias, 0, nt, g);
#pragma unroll
  for (uint32_t s = 0; s < 6; ++s) {
    const bf16x8 bop = encode(x[s], g);
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) store_col(xt, ld, g_xt_row0(0) + g_kperm(s, g, j), b, bop[j]);
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt) acc[nt] = mfma(wf[(nt * 6 + s) * 64 + lane], bop, acc[nt]);
  }
#pragma unroll
  for (int l = 1; l < PT_G_LAYERS; ++l) {
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (acc[nt][j] > 0.f) mask[l - 1] |= 1u << (4 * nt + j);
    bf16x8 in[2] = {to_operand<true>(acc[0], acc[1]), to_operand<true>(acc[2], acc[3])};
#pragma unroll
    for (uint32_t s = 0; s < 2; ++s)
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) store_col(xt, ld, g_xt_row0(l) + g_kperm(s, g, j), b, in[s][j]);
    const bf16x8 *wl = wf + g_fwd_frag_offset(l) / 8;
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt) acc[nt] = bias_tile(bias, l, nt, g);
#pragma unroll
    for (uint32_t s = 0; s < 2; ++s)
#pragma unroll
      for (uint32_t nt = 0; nt < 4; ++nt) acc[nt] = mfma(wl[(nt * 2 + s) * 64 + lane], in[s], acc[nt]);
  }
  // -- loss -w (log p_g + log(c p_g + (1 - c) p_bsdf)) and its gradient at the raw outputs.
  // The first term keeps the lobes learning the radiance while c is small, the
  // second trains c (and the lobes) for the mixture.  Gradients go through the
  // lobe responsibilities r_i = pi_i v_i / p_g (log-sum-exp), so no 1/p_g appears:
  //   d(-log p_g - log p_mix) = -(1 + rho) d log p_g - (p_g - p_bsdf) / p_mix dc,
  //   rho = c p_g / p_mix in [0, 1].
  LobeQuad r;
  g_process(acc, g, gv.c_lo, gv.c_span, r);
  float lj[4], dot_mw[4], dlv_dk[4];
  float lmax = -INFINITY;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    dot_mw[j] = r.mux[j] * wx + r.muy[j] * wy + r.muz[j] * wz;
    const float k = r.kappa[j];
    const float e2 = expf(-2.f * k);
    lj[j] = logf(r.pi[j]) + g_log_norm(k) + k * (dot_mw[j] - 1.f);  // log(pi_i v_i)
    dlv_dk[j] = 1.f / k - 2.f * e2 / (1.f - e2) + dot_mw[j] - 1.f;  // d log v_i / d kappa
    lmax = fmaxf(lmax, lj[j]);
  }
  if (g >= 2u) lmax = -INFINITY;
  lmax = fmaxf(lmax, __shfl_xor(lmax, 16));  // groups 0 and 1 hold the 8 lobes
  float esum = 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) esum += (g < 2u) ? expf(lj[j] - lmax) : 0.f;
  esum += __shfl_xor(esum, 16);
  float log_pg = lmax + logf(esum);
  log_pg = __shfl(log_pg, (int)q);  // every lane of the point (group 2 needs it for c)
  const float c = __shfl(r.c, (int)(q | 32u));
  const float sg = __shfl(r.s_sig, (int)(q | 32u));
  const float pg = expf(log_pg);
  const float pm = c * pg + (1.f - c) * pb;
  const float log_pm = pb > 0.f ? logf(pm) : logf(c) + log_pg;
  const float rho = expf(logf(c) + log_pg - log_pm);
  const bool live = b < batch && wt > 0.f && isfinite(log_pg) && isfinite(log_pm);
  const float coef = live ? -wt * (1.f + rho) * inv_batch : 0.f;  // times d log p_g
  float loss = (live && g == 0u) ? -wt * (log_pg + log_pm) * inv_batch : 0.f;
  f32x4 dout[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) dout[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  float dmuy[4] = {0.f, 0.f, 0.f, 0.f};
  if (g < 2u) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float resp = expf(lj[j] - log_pg);
      dout[0][j] = coef * (resp - r.pi[j]);                                      // logits
      dout[1][j] = coef * resp * dlv_dk[j] * g_dkappa(acc[1][j], r.kappa[j]);    // concentration
      const float gm = coef * resp * r.kappa[j] * r.inv_len[j];
      dout[2][j] = gm * (wx - dot_mw[j] * r.mux[j]);
      dmuy[j] = gm * (wy - dot_mw[j] * r.muy[j]);
      dout[3][j] = gm * (wz - dot_mw[j] * r.muz[j]);
    }
  }
  // y components go back to the lanes of group g + 2 (tile 2, rows 40 + 4 g' + j)
  float back[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) back[j] = __shfl_xor(dmuy[j], 32);
  if (g >= 2u) {
#pragma unroll
    for (int j = 0; j < 4; ++j) dout[2][j] = back[j];
#pragma unroll
    for (int j = 0; j < 4; ++j) dout[3][j] = 0.f;  // rows 56..63 carry no output
#pragma unroll
    for (int j = 0; j < 4; ++j) dout[1][j] = 0.f;  // rows 24..31
    if (g == 2u) {
      // (p_g - p_bsdf) / p_mix = rho / c - (1 - rho) / (1 - c)
      const float dsel = live ? -wt * inv_batch * (rho / c - (1.f - rho) / (1.f - c)) : 0.f;
      dout[0][0] = dsel * gv.c_span * sg * (1.f - sg);  // selection logit (row 8)
      dout[0][1] = dout[0][2] = dout[0][3] = 0.f;
    } else {
      dout[0] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  }
  // -- backward
  f32x4 dz[4] = {dout[0], dout[1], dout[2], dout[3]};
  for (int l = PT_G_LAYERS - 1; l >= 0; --l) {
    // bias gradient partial and the transposed output gradient of layer l
    float *dbp = dbpart + ((size_t)wave * PT_G_LAYERS + l) * 64u;
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt) {
      const f32x4 cs = tile_colsum(dz[nt]);
      if (q == 0u) {
#pragma unroll
        for (int j = 0; j < 4; ++j) dbp[16 * nt + 4 * g + j] = cs[j];
      }
    }
    const bf16x8 dzb[2] = {to_operand<false>(dz[0], dz[1]), to_operand<false>(dz[2], dz[3])};
#pragma unroll
    for (uint32_t s = 0; s < 2; ++s)
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) store_col(dzt, ld, 64u * l + g_kperm(s, g, j), b, dzb[s][j]);
    if (l == 0) break;
    // dX_l^T = W_l dZ_l^T, then the ReLU of layer l - 1
    const bf16x8 *wl = wb + g_bwd_frag_offset(l) / 8;
    f32x4 dx[4];
#pragma unroll
    for (uint32_t kt = 0; kt < 4; ++kt) dx[kt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (uint32_t s = 0; s < 2; ++s)
#pragma unroll
      for (uint32_t kt = 0; kt < 4; ++kt) dx[kt] = mfma(wl[(kt * 2 + s) * 64 + lane], dzb[s], dx[kt]);
#pragma unroll
    for (uint32_t kt = 0; kt < 4; ++kt)
#pragma unroll
      for (int j = 0; j < 4; ++j) dz[kt][j] = ((mask[l - 1] >> (4 * kt + j)) & 1u) ? dx[kt][j] : 0.f;
  }
  // loss partial of the wave
  for (int off = 1; off < 64; off <<= 1) loss += __shfl_xor(loss, off);
  if (lane == 0u) losspart[wave] = loss;
}

// Weight gradients: wave (tile, split) sums dW_l^T[16 nt + m][16 kt + k] over
// its share of the batch columns.  Tiles: layer 0 4 x 12, layers 1..4 4 x 4.
PT_HD void g_tile(uint32_t t, int &l, uint32_t &nt, uint32_t &kt) {
  if (t < 48) {
    l = 0;
    nt = t / 12;
    kt = t % 12;
  } else {
    t -= 48;
    l = 1 + (int)(t / 16);
    nt = (t % 16) / 4;
    kt = t % 4;
  }
}
#define PT_G_TILES 112
__global__ void __launch_bounds__(64) k_guide_train_b(const uint16_t *xt, const uint16_t *dzt, uint32_t ld,
                                                      uint32_t splits, float *gpart) {
  const uint32_t tile = blockIdx.x % PT_G_TILES, split = blockIdx.x / PT_G_TILES;
  int l;
  uint32_t nt, kt;
  g_tile(tile, l, nt, kt);
  const uint32_t lane = threadIdx.x, m = lane & 15u, g = lane >> 4;
  const uint32_t steps = ld / 32u, per = (steps + splits - 1) / splits;
  const uint32_t s0 = split * per, s1 = min(steps, s0 + per);
  const uint16_t *arow = dzt + (size_t)(64u * l + 16u * nt + m) * ld + 8u * g;
  const uint16_t *brow = xt + (size_t)(g_xt_row0(l) + 16u * kt + m) * ld + 8u * g;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (uint32_t s = s0; s < s1; ++s) {
    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(arow + 32u * s);
    const bf16x8 bb = *reinterpret_cast<const bf16x8 *>(brow + 32u * s);
    acc = mfma(a, bb, acc);
  }
  // C[4 g + j][m]: dW_l[k = 16 kt + m][row = 16 nt + 4 g + j]
  float *gp = gpart + (size_t)split * PT_G_NPARAMS;
  const uint32_t k = 16u * kt + m, N = g_layer_n(l);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t row = 16u * nt + 4u * g + j;
    int o = -1;
    if (l < 4) {
      o = (int)row;
    } else {
      for (uint32_t oo = 0; oo < PT_G_OUT; ++oo)
        if (g_out_row(oo) == row) o = (int)oo;
    }
    if (o >= 0) gp[g_w_offset(l) + k * N + (uint32_t)o] = acc[j];
  }
}

// Position of parameter p: layer, bias or weight, input k, output o.
PT_HD void g_param(uint32_t p, int &l, bool &is_bias, uint32_t &k, uint32_t &o) {
  l = 0;
  while (l < 4 && p >= g_w_offset(l + 1)) ++l;
  uint32_t r = p - g_w_offset(l);
  const uint32_t K = g_layer_k(l), N = g_layer_n(l);
  if (r < K * N) {
    is_bias = false;
    k = r / N;
    o = r % N;
  } else {
    is_bias = true;
    k = 0;
    o = r - K * N;
  }
}
PT_HD uint32_t g_slot_of(uint32_t r32, uint32_t &g, uint32_t &j) {  // inverse of g_kperm within a k-step
  if (r32 < 16u) {
    g = r32 / 4u;
    j = r32 % 4u;
  } else {
    g = (r32 - 16u) / 4u;
    j = 4u + (r32 - 16u) % 4u;
  }
  return 0u;
}

// Bias gradients: one wave per bias, summing its per-tile partials.
__global__ void __launch_bounds__(64) k_guide_bias_sum(const float *dbpart, uint32_t nwaves, float *bsum) {
  const uint32_t lb = blockIdx.x;  // layer * 64 + row
  const uint32_t l = lb / 64u, row = lb % 64u;
  float acc = 0.f;
  for (uint32_t wv = threadIdx.x; wv < nwaves; wv += 64u) acc += dbpart[((size_t)wv * PT_G_LAYERS + l) * 64u + row];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (threadIdx.x == 0) bsum[lb] = acc;
}

// mode 0: pack the master weights; mode 1: Adam step (gradient = sum of the
// partials; biases from k_guide_bias_sum) and pack; mode 2: the gradient
// alone, into mom.
__global__ void k_guide_update(float *params, float *mom, float *vel, const float *gpart, uint32_t splits,
                               const float *dbpart, uint32_t nwaves, float lr, float b1, float b2, float eps,
                               float c1, float c2, int mode, uint16_t *fwd, uint16_t *bwd, float *bias) {
  const uint32_t p = g_tid();
  if (p >= PT_G_NPARAMS) return;
  int l;
  bool is_bias;
  uint32_t k, o;
  g_param(p, l, is_bias, k, o);
  const uint32_t row = l == 4 ? g_out_row(o) : o;
  float th = params[p];
  if (mode >= 1) {
    float gsum = 0.f;
    if (is_bias) {
      gsum = dbpart[l * 64 + row];  // the reduced sums
      (void)nwaves;
    } else {
      for (uint32_t s = 0; s < splits; ++s) gsum += gpart[(size_t)s * PT_G_NPARAMS + p];
    }
    if (mode == 2) {
      mom[p] = gsum;
      return;
    }
    const float m1 = b1 * mom[p] + (1.f - b1) * gsum;
    const float v1 = b2 * vel[p] + (1.f - b2) * gsum * gsum;
    mom[p] = m1;
    vel[p] = v1;
    th -= lr * (m1 / c1) / (sqrtf(v1 / c2) + eps);
    params[p] = th;
  }
  if (is_bias) {
    bias[l * 64 + row] = th;
    return;
  }
  const __bf16 hv = (__bf16)th;
  const uint16_t bits = __builtin_bit_cast(uint16_t, hv);
  {  // forward fragment F_l[nt][s]: lane (row % 16) + 16 g, slot j of input k
    const uint32_t KS = g_layer_k(l) / 32u;
    uint32_t g, j;
    g_slot_of(k % 32u, g, j);
    const uint32_t lane = (row % 16u) + 16u * g;
    fwd[g_fwd_frag_offset(l) + (((row / 16u) * KS + k / 32u) * 64u + lane) * 8u + j] = bits;
  }
  if (l >= 1) {  // backward fragment G_l[kt][s]: lane (k % 16) + 16 g, slot j of row
    uint32_t g, j;
    g_slot_of(row % 32u, g, j);
    const uint32_t lane = (k % 16u) + 16u * g;
    bwd[g_bwd_frag_offset(l) + (((k / 16u) * 2u + row / 32u) * 64u + lane) * 8u + j] = bits;
  }
}

// Training samples from the path records of a pass: record (slot / stride,
// vertex) holds the point's inputs, the sampled direction, p_bsdf, the
// throughput beta at the vertex and the slot's radiance once the next vertex
// was reached; F / q = (L_final - L_after) / beta.
__global__ void k_guide_collect(Wave w, uint32_t n_rec, uint32_t verts, uint32_t stride, uint32_t pass,
                                float *samples, uint32_t cap, uint32_t *count) {
  const uint32_t r = g_tid();
  bool keep = false;
  float s = 0.f;
  const float *rec = w.grec + (size_t)min(r, n_rec * verts - 1u) * PT_G_REC;
  if (r < n_rec * verts && __float_as_uint(rec[16]) == pass && __float_as_uint(rec[17]) == pass) {
    const uint32_t slot = (r / verts) * stride;
    const f3 lf = mk3(w.lx[slot], w.ly[slot], w.lz[slot]);
  // the next vertex's emission at full weight: the radiance along the direction
  const f3 la = mk3(rec[13] - rec[18], rec[14] - rec[19], rec[15] - rec[20]);
    const f3 bt = mk3(rec[10], rec[11], rec[12]);
    s += bt.x > 0.f ? (lf.x - la.x) / bt.x : 0.f;
    s += bt.y > 0.f ? (lf.y - la.y) / bt.y : 0.f;
    s += bt.z > 0.f ? (lf.z - la.z) / bt.z : 0.f;
    s *= (1.f / 3.f);
    keep = s > 0.f && isfinite(s) && rec[9] >= 0.f;
  }
  // converged here: one atomic per wave
  const uint32_t idx = wave_append(count, keep);
  if (!keep || idx >= cap) return;
  float *o = samples + (size_t)idx * PT_G_SAMPLE;
  for (int v = 0; v < 10; ++v) o[v] = rec[v];
  o[10] = s;
  o[11] = 0.f;
}

// ------------------------------------------------------------------ launchers
static uint32_t blocks_of(size_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

void launch_guide_infer(hipStream_t st, const DevScene &s, const Wave &w, int qi, const GuideView &gv,
                        uint32_t kinds, uint32_t grid, bool per_pixel, uint64_t chunk_first, uint32_t sppc,
                        uint32_t n) {
  InferPixels px{chunk_first, sppc, n, per_pixel && n > 0u && sppc > 0u ? 1u : 0u};
  hipLaunchKernelGGL(k_guide_infer, dim3(grid), dim3(PT_GI_BLOCK), 0, st, s, w, qi, gv, kinds, px);
}
void launch_guide_points(hipStream_t st, const GuideView &gv, const float *xs, size_t n, float *raw, float *mix) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_guide_points, dim3(blocks_of(n, 256)), dim3(256), 0, st, gv, xs, n, raw, mix);
}
void launch_guide_mix(hipStream_t st, const float *mix, const float *dirs, const float *u, size_t n, float *pdf,
                      float *out_dir, float *out_pdf) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_guide_mix, dim3(blocks_of(n, 256)), dim3(256), 0, st, mix, dirs, u, n, pdf, out_dir,
                     out_pdf);
}
void launch_guide_train(hipStream_t st, const GuideView &gv, const float *samples, uint32_t n_samples,
                        uint32_t batch, uint32_t pick_seed, const GuideTrainBufs &tb) {
  const uint32_t ld = tb.ld;  // batch rounded up to 64
  hipLaunchKernelGGL(k_guide_train_a, dim3(ld / 64u), dim3(256), 0, st, gv, samples, n_samples, batch, ld,
                     pick_seed, tb.xt, tb.dzt, tb.dbpart, tb.losspart);
  hipLaunchKernelGGL(k_guide_train_b, dim3(PT_G_TILES * tb.splits), dim3(64), 0, st, tb.xt, tb.dzt, ld,
                     tb.splits, tb.gpart);
}
void launch_guide_update(hipStream_t st, const GuideParamBufs &pb, const GuideTrainBufs *tb, float lr, float b1,
                         float b2, float eps, float c1, float c2, int mode) {
  if (tb) hipLaunchKernelGGL(k_guide_bias_sum, dim3(PT_G_LAYERS * 64), dim3(64), 0, st, tb->dbpart, tb->ld / 16u,
                             tb->bsum);
  hipLaunchKernelGGL(k_guide_update, dim3(blocks_of(PT_G_NPARAMS, 256)), dim3(256), 0, st, pb.params, pb.mom,
                     pb.vel, tb ? tb->gpart : nullptr, tb ? tb->splits : 0u, tb ? tb->bsum : nullptr,
                     tb ? tb->ld / 16u : 0u, lr, b1, b2, eps, c1, c2, mode, pb.fwd, pb.bwd, pb.bias);
}
void launch_guide_collect(hipStream_t st, const Wave &w, uint32_t n_rec, uint32_t verts, uint32_t stride,
                          uint32_t pass, float *samples, uint32_t cap, uint32_t *count) {
  const size_t n = (size_t)n_rec * verts;
  if (n == 0) return;
  hipLaunchKernelGGL(k_guide_collect, dim3(blocks_of(n, 256)), dim3(256), 0, st, w, n_rec, verts, stride, pass,
                     samples, cap, count);
}

}  // namespace pt
