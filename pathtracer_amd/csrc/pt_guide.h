// pt_guide.h -- the learned path-guiding proposal: a small MLP that maps a
// surface point (position, normal) to a mixture of 8 von Mises-Fisher lobes
// over incident directions plus the probability of sampling that mixture
// instead of the BSDF.  Shared by the kernels (pt_guide.hip, pt_kernels.hip)
// and the host (pt_guide.cpp).  There is no counterpart in the reference.
//
// Network (BF16 weights and activations, FP32 accumulation and biases):
//   input   6 values in [0,1]: the position in the scene's bounding box and
//           the normal mapped by (n + 1) / 2
//   enc     one-blob encoding, 32 bins per value: feature (v, b) =
//           bump(b + 0.5 - 32 x_v), bump(t) = max(0, 1 - t^2 / 9)^2   (192)
//   L0..L3  dense 192->64, 64->64, 64->64, 64->64, each + bias, ReLU
//   L4      dense 64->41 + bias (raw outputs)
// Raw output o (41 values):
//   o[0..7]   lobe logits a_i        -> pi = softmax(a)
//   o[8]      selection logit s      -> c = c_lo + c_span sigmoid(s), in a range
//                                       set per guide (default [0.05, 0.5])
//   o[9..16]  concentration r_i      -> kappa_i = exp(1.5 + 6.5 tanh(r_i / 6.5)),
//                                       smoothly within [e^-5, e^8]
//   o[17..40] mean directions, x of lobes 0..7, then y, then z
//             -> mu_i = normalize(o[17+i], o[25+i], o[33+i])
// Guide pdf: p_g(w) = sum_i pi_i kappa_i / (2 pi (1 - e^{-2 kappa_i}))
//                               * exp(kappa_i (mu_i . w - 1)).
// Sampling: a direction is drawn from the guide with probability c, else
// from the BSDF; both are weighted by the one-sample MIS mixture pdf
// c p_g + (1 - c) p_bsdf, so the estimator stays unbiased for any weights.
//
// Parameters (FP32 master copy, the order of pt_guide_get_params): for layer
// l = 0..4, W_l[k][n] row-major (k input, n output), then b_l[n].
#pragma once
#include <stdint.h>
#include "pt_math.h"
#include "../../include/pathtracer_amd.h"

namespace pt {

#define PT_G_IN 6
#define PT_G_BINS 32
#define PT_G_K0 (PT_G_IN * PT_G_BINS)  // 192 encoded features
#define PT_G_H 64                       // hidden width
#define PT_G_OUT 41                     // raw outputs
#define PT_G_LOBES 8
#define PT_G_LAYERS 5
// Parameter counts and offsets in the flat FP32 order.
#define PT_G_NPARAMS (PT_G_K0 * PT_G_H + PT_G_H + 3 * (PT_G_H * PT_G_H + PT_G_H) + PT_G_H * PT_G_OUT + PT_G_OUT)

PT_HD uint32_t g_layer_k(int l) { return l == 0 ? PT_G_K0 : PT_G_H; }
PT_HD uint32_t g_layer_n(int l) { return l == 4 ? PT_G_OUT : PT_G_H; }
PT_HD uint32_t g_w_offset(int l) {
  uint32_t o = 0;
  for (int i = 0; i < l; ++i) o += g_layer_k(i) * g_layer_n(i) + g_layer_n(i);
  return o;
}
PT_HD uint32_t g_b_offset(int l) { return g_w_offset(l) + g_layer_k(l) * g_layer_n(l); }

// On the device the output layer is padded to 64 rows; raw output o sits in
// row g_out_row(o), chosen so that after the last product each lane of a
// 16-query tile holds the logits, concentrations and x, z mean components of
// four lobes of one query (and the y components sit 32 lanes away).
PT_HD uint32_t g_out_row(uint32_t o) { return o < 9 ? o : (o < 17 ? 16 + (o - 9) : 32 + (o - 17)); }

// The MFMA operand layout (v_mfma_f32_16x16x32_bf16, checked on the device):
// lane l holds A[m = l % 16][k = 8 (l / 16) + j] and B[k = 8 (l / 16) + j][n =
// l % 16], j = 0..7, and receives C[m = 4 (l / 16) + j][n = l % 16], j = 0..3.
// Layers are computed transposed, Y^T = W^T X^T (m = output unit, n =
// query), so the C tiles of one layer are the B operand of the next once the
// 32 inputs of k-step s are ordered as: slot (g, j) of k-step s is input
// 32 s + (j < 4 ? 4 g + j : 16 + 4 g + j - 4).
PT_HD uint32_t g_kperm(uint32_t s, uint32_t g, uint32_t j) {
  return 32u * s + (j < 4u ? 4u * g + j : 16u + 4u * g + (j - 4u));
}

// Packed BF16 weights on the device.  Forward fragments F_l[nt][s] (64 lanes
// x 8 values) hold A = W_l^T: lane l, j -> W_l[g_kperm(s, l/16, j)][16 nt + l%16];
// backward fragments G_l[kt][s] (l = 1..4) hold A = W_l: lane l, j ->
// W_l[16 kt + l%16][row g_kperm(s, l/16, j)].  Rows of the padded output
// layer that carry no raw output are zero.
#define PT_G_FRAG 512  // bf16 values per fragment
PT_HD uint32_t g_fwd_frag_offset(int l) {  // in bf16 values
  return l == 0 ? 0u : (uint32_t)(4 * 6 + (l - 1) * 4 * 2) * PT_G_FRAG;
}
#define PT_G_FWD_VALUES ((4 * 6 + 4 * 4 * 2) * PT_G_FRAG)  // 28672
PT_HD uint32_t g_bwd_frag_offset(int l) { return (uint32_t)((l - 1) * 4 * 2) * PT_G_FRAG; }  // l = 1..4
#define PT_G_BWD_VALUES (4 * 4 * 2 * PT_G_FRAG)  // 16384
#define PT_G_BIAS_VALUES (5 * 64)  // FP32, padded rows (output layer by g_out_row) zero

// Processed mixture of one shading point as the renderer reads it (48 floats):
//   [4 i .. 4 i + 3]  (kappa_i mu_i, log(pi_i kappa_i / (2 pi (1 - e^{-2 kappa_i}))) - kappa_i)
//   [32 + i]          pi_i
//   [40]              c (0: this point is not guided)
#define PT_G_MIX 48
// Training record of one guided vertex (24 floats): [0..5] the point's
// inputs, [6..8] the sampled direction, [9] p_bsdf of it, [10..12] the
// throughput arriving at the vertex, [13..15] the slot's radiance when the
// path reached the next vertex, [16] / [17] the pass that wrote the vertex
// fields / the radiance (as uint32 bits), [18..20] the share of the next
// vertex's emission that MIS gave to light sampling (the training target is
// the radiance arriving along the direction, unweighted).
#define PT_G_REC 24
// kappa from the raw concentration output, and d kappa / d r.
PT_HD float g_kappa(float r) { return expf(1.5f + 6.5f * tanhf(r * (1.f / 6.5f))); }
PT_HD float g_dkappa(float r, float kappa) {
  const float t = tanhf(r * (1.f / 6.5f));
  return kappa * (1.f - t * t);
}

// Training scratch, transposed (bf16, row stride = the batch padded to 64):
// layer inputs X_0 (192 rows) and X_1..X_4 (64 rows each) from row
// g_xt_row0(l); output gradients dZ_0..dZ_4 (64 rows each, the output layer
// in padded rows) from row 64 l.
PT_HD uint32_t g_xt_row0(int l) { return l == 0 ? 0u : 192u + 64u * (uint32_t)(l - 1); }
#define PT_G_XT_ROWS 448
#define PT_G_DZT_ROWS 320

struct GuideView {
  const uint16_t *fwd;   // forward fragments (bf16 bits)
  const uint16_t *bwd;   // backward fragments
  const float *bias;     // 5 x 64
  float lo[3], inv_ext[3];  // position normalisation: (p - lo) * inv_ext
  float c_lo, c_span;       // selection probability range
};

// Materials the guide samples for: kinds in the mask whose lobes are all
// rough (a delta lobe takes no guided directions).  Conductor alphas are
// p[6], p[7]; the Disney roughness is p[7].
PT_HD bool g_guidable(uint32_t kind, const float *p, uint32_t kinds) {
  if (!((kinds >> kind) & 1u)) return false;
  if (kind == PT_MAT_LAMBERTIAN) return true;
  if (kind == PT_MAT_CONDUCTOR) return fminf(p[6], p[7]) >= 1e-3f;
  if (kind == PT_MAT_DISNEY) return p[7] >= 0.05f;
  return false;
}

// One-blob bump of the encoding.
PT_HD float g_bump(float t) {
  float u = fmaxf(0.f, 1.f - t * t * (1.f / 9.f));
  return u * u;
}

// vMF helpers on the processed mixture.
PT_HD float g_log_norm(float kappa) {  // log(kappa / (2 pi (1 - e^{-2 kappa})))
  return logf(kappa) - logf(2.f * PT_PI) - log1pf(-expf(-2.f * kappa));
}
PT_D float g_pdf(const float *mix, f3 w) {
  float p = 0.f;
#pragma unroll
  for (int i = 0; i < PT_G_LOBES; ++i)
    p += expf(mix[4 * i] * w.x + mix[4 * i + 1] * w.y + mix[4 * i + 2] * w.z + mix[4 * i + 3]);
  return p;
}
// Samples lobe i by pi with u0, then a direction about mu_i with (u1, u2):
// cos(theta) = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa, phi = 2 pi u2.
PT_D f3 g_sample(const float *mix, float u0, float u1, float u2) {
  int i = 0;
  float acc = mix[32];
  while (i < PT_G_LOBES - 1 && u0 >= acc) {
    ++i;
    acc += mix[32 + i];
  }
  f3 m = mk3(mix[4 * i], mix[4 * i + 1], mix[4 * i + 2]);
  float kappa = len3(m);
  f3 mu = divs3(m, kappa);
  float ct = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * kappa)) / kappa;
  ct = fminf(1.f, fmaxf(-1.f, ct));
  float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
  float ph = 2.f * PT_PI * u2;
  // any orthonormal frame about mu (Duff et al. 2017)
  float sg = copysignf(1.f, mu.z);
  float a = -1.f / (sg + mu.z), b = mu.x * mu.y * a;
  f3 t1 = mk3(1.f + sg * mu.x * mu.x * a, sg * b, -sg * mu.x);
  f3 t2 = mk3(b, sg + mu.y * mu.y * a, -mu.y);
  return add3(add3(scl3(t1, st * cosf(ph)), scl3(t2, st * sinf(ph))), scl3(mu, ct));
}

}  // namespace pt
