// pt_guide.cpp -- host side of the learned path-guiding proposal: parameter
// initialisation and packing, Adam steps, and the C ABI of the guide
// (include/pathtracer_amd.h).  No counterpart in the reference.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pt_guide_host.h"

using namespace pt;

static const float kB1 = 0.9f, kB2 = 0.999f, kEps = 1e-8f;

// Seeded initial weights: He-uniform hidden layers; a small output layer whose
// biases spread the 8 lobes over the cube's diagonals (kappa = 2, c = 0.5).
static void init_params(uint32_t seed, std::vector<float> &p) {
  p.assign(PT_G_NPARAMS, 0.f);
  uint64_t st = 0x853c49e6748fea9bull ^ ((uint64_t)seed << 1);
  auto next = [&]() {  // PCG-XSH-RR, uniform in [0, 1)
    uint64_t old = st;
    st = old * 6364136223846793005ull + 1442695040888963407ull;
    uint32_t x = (uint32_t)(((old >> 18u) ^ old) >> 27u), r = (uint32_t)(old >> 59u);
    return (float)(((x >> r) | (x << ((32u - r) & 31u))) >> 8) * (1.f / 16777216.f);
  };
  for (int l = 0; l < PT_G_LAYERS; ++l) {
    const uint32_t K = g_layer_k(l), N = g_layer_n(l);
    const float a = std::sqrt(6.f / (float)K) * (l == 4 ? 0.1f : 1.f);
    for (uint32_t i = 0; i < K * N; ++i) p[g_w_offset(l) + i] = (2.f * next() - 1.f) * a;
  }
  float *b4 = &p[g_b_offset(4)];
  for (int i = 0; i < PT_G_LOBES; ++i) {
    b4[9 + i] = 6.5f * std::atanh((std::log(2.f) - 1.5f) / 6.5f);  // kappa = 2
    b4[17 + i] = (i & 1) ? 1.f : -1.f;
    b4[25 + i] = (i & 2) ? 1.f : -1.f;
    b4[33 + i] = (i & 4) ? 1.f : -1.f;
  }
}

static pt_status upload_params(pt_guide *g, const float *p) {
  HIP_TRY(hipMemcpy(g->params.p, p, PT_G_NPARAMS * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(g->mom.p, 0, PT_G_NPARAMS * sizeof(float)));
  HIP_TRY(hipMemset(g->vel.p, 0, PT_G_NPARAMS * sizeof(float)));
  g->adam_t = 0;
  launch_guide_update(g->stream, g->pbufs(), nullptr, 0.f, kB1, kB2, kEps, 1.f, 1.f, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  return PT_OK;
}

extern "C" pt_status pt_guide_create(pt_scene *scene, uint32_t seed, pt_guide **out) {
  if (!scene || !out) return fail(PT_ERR_INVALID, "pt_guide_create: NULL argument");
  float lo[3], hi[3];
  int dev = -1;
  pt_status st = ptx_scene_box(scene, lo, hi, &dev);
  if (st != PT_OK) return st;
  pt_guide *g = new (std::nothrow) pt_guide();
  if (!g) return fail(PT_ERR_NOMEM, "pt_guide_create: out of host memory");
  g->device = dev;
  for (int a = 0; a < 3; ++a) {
    const float ext = hi[a] - lo[a];
    g->lo[a] = lo[a];
    g->inv_ext[a] = ext > 0.f ? 1.f / ext : 1.f;
  }
  auto cleanup = [&](pt_status s) {
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return s;
  };
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(PT_ERR_DEVICE, "pt_guide_create: stream creation failed"));
  const size_t pb = PT_G_NPARAMS * sizeof(float);
  if (g->params.alloc(pb) != hipSuccess || g->mom.alloc(pb) != hipSuccess || g->vel.alloc(pb) != hipSuccess ||
      g->fwd.alloc(PT_G_FWD_VALUES * 2) != hipSuccess || g->bwd.alloc(PT_G_BWD_VALUES * 2) != hipSuccess ||
      g->bias.alloc(PT_G_BIAS_VALUES * sizeof(float)) != hipSuccess)
    return cleanup(fail(PT_ERR_NOMEM, "pt_guide_create: device allocation failed"));
  // padded fragment slots and bias rows stay zero
  if (hipMemset(g->fwd.p, 0, PT_G_FWD_VALUES * 2) != hipSuccess ||
      hipMemset(g->bwd.p, 0, PT_G_BWD_VALUES * 2) != hipSuccess ||
      hipMemset(g->bias.p, 0, PT_G_BIAS_VALUES * sizeof(float)) != hipSuccess)
    return cleanup(fail(PT_ERR_DEVICE, "pt_guide_create: memset failed"));
  std::vector<float> p;
  init_params(seed, p);
  st = upload_params(g, p.data());
  if (st != PT_OK) return cleanup(st);
  *out = g;
  return PT_OK;
}

extern "C" void pt_guide_destroy(pt_guide *g) {
  if (!g) return;
  if (g->stream) {
    (void)hipStreamSynchronize(g->stream);
    (void)hipStreamDestroy(g->stream);
  }
  delete g;
}

extern "C" pt_status pt_guide_get_params(pt_guide *g, float *params) {
  if (!g || !params) return fail(PT_ERR_INVALID, "pt_guide_get_params: NULL argument");
  HIP_TRY(hipStreamSynchronize(g->stream));
  HIP_TRY(hipMemcpy(params, g->params.p, PT_G_NPARAMS * sizeof(float), hipMemcpyDeviceToHost));
  return PT_OK;
}

extern "C" pt_status pt_guide_set_params(pt_guide *g, const float *params) {
  if (!g || !params) return fail(PT_ERR_INVALID, "pt_guide_set_params: NULL argument");
  for (size_t i = 0; i < PT_G_NPARAMS; ++i)
    if (!std::isfinite(params[i])) return fail(PT_ERR_INVALID, "pt_guide_set_params: non-finite parameter");
  return upload_params(g, params);
}

extern "C" pt_status pt_guide_forward(pt_guide *g, const float *x, size_t n, float *raw, float *mix) {
  if (!g || (n && !x)) return fail(PT_ERR_INVALID, "pt_guide_forward: NULL argument");
  if (n == 0) return PT_OK;
  if (n > (1u << 26)) return fail(PT_ERR_INVALID, "pt_guide_forward: too many points");
  HIP_TRY(upload(g->io_a, x, n * PT_G_IN * sizeof(float)));
  if (raw) HIP_TRY(g->io_b.alloc(n * PT_G_OUT * sizeof(float)));
  if (mix) HIP_TRY(g->io_c.alloc(n * PT_G_MIX * sizeof(float)));
  launch_guide_points(g->stream, g->view(), g->io_a.as<float>(), n, raw ? g->io_b.as<float>() : nullptr,
                      mix ? g->io_c.as<float>() : nullptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  if (raw) HIP_TRY(hipMemcpy(raw, g->io_b.p, n * PT_G_OUT * sizeof(float), hipMemcpyDeviceToHost));
  if (mix) HIP_TRY(hipMemcpy(mix, g->io_c.p, n * PT_G_MIX * sizeof(float), hipMemcpyDeviceToHost));
  return PT_OK;
}

extern "C" pt_status pt_guide_eval(const float *mix, const float *dirs, const float *u, size_t n, float *pdf,
                                   float *sample_dir, float *sample_pdf) {
  if (n == 0) return PT_OK;
  if (!mix || (dirs && !pdf) || (u && (!sample_dir || !sample_pdf)))
    return fail(PT_ERR_INVALID, "pt_guide_eval: NULL argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PT_ERR_NO_DEVICE, "no HIP device visible");
  DevBuf dm, dd, du, dp, dsd, dsp;
  HIP_TRY(upload(dm, mix, n * PT_G_MIX * sizeof(float)));
  if (dirs) {
    HIP_TRY(upload(dd, dirs, n * 3 * sizeof(float)));
    HIP_TRY(dp.alloc(n * sizeof(float)));
  }
  if (u) {
    HIP_TRY(upload(du, u, n * 3 * sizeof(float)));
    HIP_TRY(dsd.alloc(n * 3 * sizeof(float)));
    HIP_TRY(dsp.alloc(n * sizeof(float)));
  }
  launch_guide_mix(nullptr, dm.as<float>(), dirs ? dd.as<float>() : nullptr, u ? du.as<float>() : nullptr, n,
                   dp.as<float>(), dsd.as<float>(), dsp.as<float>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  if (dirs) HIP_TRY(hipMemcpy(pdf, dp.p, n * sizeof(float), hipMemcpyDeviceToHost));
  if (u) {
    HIP_TRY(hipMemcpy(sample_dir, dsd.p, n * 3 * sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(sample_pdf, dsp.p, n * sizeof(float), hipMemcpyDeviceToHost));
  }
  return PT_OK;
}

pt_status ptx_guide_reserve(pt_guide *g, uint32_t batch) {
  const uint32_t ld = (batch + 63u) / 64u * 64u;
  if (ld <= g->ld && g->xt.p) return PT_OK;
  HIP_TRY(g->xt.alloc((size_t)PT_G_XT_ROWS * ld * 2));
  HIP_TRY(g->dzt.alloc((size_t)PT_G_DZT_ROWS * ld * 2));
  HIP_TRY(g->gpart.alloc((size_t)g->splits * PT_G_NPARAMS * sizeof(float)));
  HIP_TRY(g->dbpart.alloc((size_t)(ld / 16u) * PT_G_LAYERS * 64 * sizeof(float)));
  HIP_TRY(g->losspart.alloc((size_t)(ld / 16u) * sizeof(float)));
  HIP_TRY(g->bsum.alloc(PT_G_BIAS_VALUES * sizeof(float)));
  g->ld = ld;
  return PT_OK;
}

// The kernels see one batch of exactly `batch` samples: ld is set per step.
static pt::GuideTrainBufs step_bufs(const pt_guide *g, uint32_t batch) {
  pt::GuideTrainBufs t = g->tbufs();
  t.ld = (batch + 63u) / 64u * 64u;
  return t;
}

pt_status ptx_guide_step(pt_guide *g, hipStream_t st, const float *samples, uint32_t n, uint32_t batch,
                         uint32_t pick_seed, float lr, float *loss) {
  pt_status s = ptx_guide_reserve(g, batch);
  if (s != PT_OK) return s;
  const pt::GuideTrainBufs tb = step_bufs(g, batch);
  launch_guide_train(st, g->view(), samples, n, batch, pick_seed, tb);
  g->adam_t += 1;
  const float c1 = 1.f - std::pow(kB1, (float)g->adam_t), c2 = 1.f - std::pow(kB2, (float)g->adam_t);
  launch_guide_update(st, g->pbufs(), &tb, lr, kB1, kB2, kEps, c1, c2, 1);
  HIP_TRY(hipGetLastError());
  if (loss) {
    std::vector<float> lp(tb.ld / 16u);
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipMemcpy(lp.data(), g->losspart.p, lp.size() * sizeof(float), hipMemcpyDeviceToHost));
    double acc = 0.0;
    for (float v : lp) acc += v;
    *loss = (float)acc;
  }
  return PT_OK;
}

static pt_status upload_samples(pt_guide *g, const float *samples, size_t n) {
  std::vector<float> v(n * PT_G_SAMPLE, 0.f);
  for (size_t i = 0; i < n; ++i) std::memcpy(&v[i * PT_G_SAMPLE], &samples[i * 11], 11 * sizeof(float));
  HIP_TRY(upload(g->io_a, v.data(), v.size() * sizeof(float)));
  return PT_OK;
}

extern "C" pt_status pt_guide_gradient(pt_guide *g, const float *samples, size_t n, float *grad, float *loss) {
  if (!g || !samples || !grad) return fail(PT_ERR_INVALID, "pt_guide_gradient: NULL argument");
  if (n == 0 || n > (1u << 22)) return fail(PT_ERR_INVALID, "pt_guide_gradient: 1..2^22 samples");
  pt_status s = upload_samples(g, samples, n);
  if (s != PT_OK) return s;
  s = ptx_guide_reserve(g, (uint32_t)n);
  if (s != PT_OK) return s;
  const pt::GuideTrainBufs tb = step_bufs(g, (uint32_t)n);
  launch_guide_train(g->stream, g->view(), g->io_a.as<float>(), (uint32_t)n, (uint32_t)n, 0u, tb);
  // mode 2: the gradient sums alone, into io_b
  HIP_TRY(g->io_b.alloc(PT_G_NPARAMS * sizeof(float)));
  pt::GuideParamBufs pb = g->pbufs();
  pb.mom = g->io_b.as<float>();
  launch_guide_update(g->stream, pb, &tb, 0.f, kB1, kB2, kEps, 1.f, 1.f, 2);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  HIP_TRY(hipMemcpy(grad, g->io_b.p, PT_G_NPARAMS * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) {
    std::vector<float> lp(tb.ld / 16u);
    HIP_TRY(hipMemcpy(lp.data(), g->losspart.p, lp.size() * sizeof(float), hipMemcpyDeviceToHost));
    double acc = 0.0;
    for (float v : lp) acc += v;
    *loss = (float)acc;
  }
  return PT_OK;
}

extern "C" pt_status pt_guide_train_step(pt_guide *g, const float *samples, size_t n, float lr, float *loss) {
  if (!g || !samples) return fail(PT_ERR_INVALID, "pt_guide_train_step: NULL argument");
  if (n == 0 || n > (1u << 22)) return fail(PT_ERR_INVALID, "pt_guide_train_step: 1..2^22 samples");
  if (!(lr > 0.f) || !std::isfinite(lr)) return fail(PT_ERR_INVALID, "pt_guide_train_step: bad learning rate");
  pt_status s = upload_samples(g, samples, n);
  if (s != PT_OK) return s;
  s = ptx_guide_step(g, g->stream, g->io_a.as<float>(), (uint32_t)n, (uint32_t)n, 0u, lr, loss);
  if (s != PT_OK) return s;
  HIP_TRY(hipStreamSynchronize(g->stream));
  return PT_OK;
}

extern "C" pt_status pt_guide_samples(pt_guide *g, float *out, size_t cap, size_t *n) {
  if (!g || !n || (cap && !out)) return fail(PT_ERR_INVALID, "pt_guide_samples: NULL argument");
  *n = 0;
  if (!g->tcount.p) return PT_OK;
  uint32_t cnt = 0;
  HIP_TRY(hipMemcpy(&cnt, g->tcount.p, sizeof(cnt), hipMemcpyDeviceToHost));
  const size_t m = std::min<size_t>(std::min<size_t>(cnt, g->tcap), cap);
  std::vector<float> v(m * PT_G_SAMPLE);
  if (m) HIP_TRY(hipMemcpy(v.data(), g->tset.p, v.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < m; ++i) std::memcpy(&out[i * 11], &v[i * PT_G_SAMPLE], 11 * sizeof(float));
  *n = std::min<size_t>(cnt, g->tcap);
  return PT_OK;
}

extern "C" pt_status pt_guide_set_select_range(pt_guide *g, float c_min, float c_max) {
  if (!g) return fail(PT_ERR_INVALID, "pt_guide_set_select_range: NULL guide");
  if (!(c_min > 0.f) || !(c_max < 1.f) || !(c_min <= c_max))
    return fail(PT_ERR_INVALID, "pt_guide_set_select_range: need 0 < c_min <= c_max < 1");
  g->c_lo = c_min;
  g->c_span = c_max - c_min;
  return PT_OK;
}
