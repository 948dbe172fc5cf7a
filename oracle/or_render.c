/*
 * or_render.c -- oracle restatement of the reference's render driver and
 * NEE integrator, plus the NEE+MIS estimator the GPU path also offers.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   SceneInfo::render      render/renderManager.hpp:51-158
 *   getRayColor (NEE)      integrator/maxdepth_nee.hpp:9-256
 *   PerspectiveCamera      base/camera.cpp:7-56
 *   FrameBuffer offsets    base/framebuffer.cpp:131-145, utils/random.hpp:47-63
 *
 * Random numbers the reference draws from gl::rand_num (lens, light choice,
 * NEE jitter, Phong lobe) come from a counter-based hash keyed by (global
 * sample index, vertex slot, seed), the same stream the GPU path uses, so a
 * path sample is reproducible on both sides.  The Halton sample index of
 * path sample k of pixel (i, j) is 1 + k + spp*(j + H*i): the order of the
 * reference's single-threaded loops (renderManager.hpp:99-115).
 *
 * Pinning (tests/test_oracle_render.py): white-furnace closed form (an
 * enclosing Lambertian sphere of albedo a lit by a uniform environment
 * converges to the geometric series), a diffuse plane under a quad light
 * against the analytic point-to-rectangle form factor, and estimator
 * agreement (reference NEE vs NEE+MIS) where the reference is unbiased.
 */
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "oracle.h"
#include "or_scene.h"

/* Materials whose BSDF only reflects (their f is not zero below the surface in
 * the reference: Lambertian::f, material.hpp:128-133). */
static int reflect_only(uint32_t kind) {
  return kind == PT_MAT_LAMBERTIAN || kind == PT_MAT_CONDUCTOR || kind == PT_MAT_PHONG;
}

enum { SLOT_LENS_X = 0, SLOT_LENS_Y = 1, SLOT_TIME = 2, SLOT_AUX = 4, SLOT_RR = 5,
       SLOT_MIS_LIGHT = 6, SLOT_MIX = 7, SLOT_NEE_OFF = 8, SLOT_RIS = 16, SLOT_NEE_LIGHT = 40,
       SLOT_GUIDE = 44 };
static inline uint32_t slot(uint32_t d, uint32_t purpose) { return d * 64u + purpose; }

/* ---- the guided estimator's mixture (no reference counterpart) ----
 * Pinned by mathematics (tests/test_oracle_guide.py): the pdf integrates to 1
 * over the sphere and equals the numpy guide oracle's mixture pdf; samples
 * reproduce each lobe's closed-form mean cosine coth(kappa) - 1/kappa and the
 * lobe weights. */
int or_scene_set_guide(or_scene *s, const float *mix, uint32_t kinds, uint32_t verts) {
  if (!s || !mix) return -1;
  memcpy(s->gmix, mix, sizeof(s->gmix));
  s->gkinds = kinds;
  s->gverts = verts;
  return 0;
}

/* Kinds in the mask whose lobes are all rough: a delta lobe takes no guided
 * direction (conductor alphas p[6], p[7]; Disney roughness p[7]). */
int or_guidable(const pt_material_desc *m, uint32_t kinds) {
  if (!((kinds >> m->kind) & 1u)) return 0;
  if (m->kind == PT_MAT_LAMBERTIAN) return 1;
  if (m->kind == PT_MAT_CONDUCTOR) return fminf(m->p[6], m->p[7]) >= 1e-3f;
  if (m->kind == PT_MAT_DISNEY) return m->p[7] >= 0.05f;
  return 0;
}

/* p(w) = sum_i pi_i kappa_i / (2 pi (1 - e^{-2 kappa_i})) e^{kappa_i (mu_i.w - 1)},
 * each term as exp(kappa_i mu_i . w + b_i) with b_i the lobe's stored log factor. */
float or_guide_pdf(const float *mix, v3 w) {
  float p = 0.f;
  for (int i = 0; i < 8; ++i) {
    const float *l = mix + 4 * i;
    p += expf(l[0] * w.x + l[1] * w.y + l[2] * w.z + l[3]);
  }
  return p;
}

/* Lobe i by the weights' running sum against u0; then about mu_i the cosine
 * 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa (the inverse of the vMF's
 * cosine CDF, Jakob 2012) and azimuth 2 pi u2, in the branchless orthonormal
 * frame of Duff et al. 2017. */
v3 or_guide_sample(const float *mix, float u0, float u1, float u2) {
  int i = 0;
  float acc = mix[32];
  while (i < 7 && u0 >= acc) acc += mix[32 + ++i];
  v3 m = V3(mix[4 * i], mix[4 * i + 1], mix[4 * i + 2]);
  float kappa = vlen(m);
  v3 mu = vdivs(m, kappa);
  float ct = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * kappa)) / kappa;
  ct = fminf(1.f, fmaxf(-1.f, ct));
  float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
  float ph = 2.f * (float)M_PI * u2;
  float sg = copysignf(1.f, mu.z);
  float a = -1.f / (sg + mu.z), b = mu.x * mu.y * a;
  v3 t1 = V3(1.f + sg * mu.x * mu.x * a, sg * b, -sg * mu.x);
  v3 t2 = V3(b, sg + mu.y * mu.y * a, -mu.y);
  return vadd(vadd(vscale(t1, st * cosf(ph)), vscale(t2, st * sinf(ph))), vscale(mu, ct));
}

int or_guide_eval_batch(const float *mix, const float *dirs, const float *u, size_t n, float *pdf,
                        float *sample_dir) {
  for (size_t i = 0; i < n; ++i) {
    const float *mx = mix + 48 * i;
    if (dirs && pdf) pdf[i] = or_guide_pdf(mx, V3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    if (u && sample_dir) {
      v3 w = or_guide_sample(mx, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
      sample_dir[3 * i] = w.x; sample_dir[3 * i + 1] = w.y; sample_dir[3 * i + 2] = w.z;
    }
  }
  return 0;
}

typedef struct {
  const or_scene *s;
  const pt_render_cfg *cfg;
  uint32_t spp, lx, ly, ns;
  float *offx, *offy;
} rctx;

static void grid_factor(uint32_t n, uint32_t *a, uint32_t *b) {
  uint32_t x = 1;
  for (uint32_t f = 1; f * f <= n; ++f)
    if (n % f == 0) x = f;
  *a = x;
  *b = n / x;
}

/* FrameBuffer::setOffsets (framebuffer.cpp:131-145): one jittered set per
 * frame, shared by every pixel. */
static void make_offsets(rctx *c) {
  uint32_t sx = c->cfg->spp_x, sy = c->cfg->spp_y;
  c->offx = (float *)malloc(sizeof(float) * c->spp);
  c->offy = (float *)malloc(sizeof(float) * c->spp);
  for (uint32_t i = 0; i < sx; ++i)
    for (uint32_t j = 0; j < sy; ++j) {
      uint32_t k = i * sy + j;
      float a0 = (float)i / (float)sx, a1 = (float)(i + 1) / (float)sx;
      float b0 = (float)j / (float)sy, b1 = (float)(j + 1) / (float)sy;
      c->offx[k] = (float)(a0 + (a1 - a0) * (double)or_rng(0xffffffffu - k, 0u, c->cfg->seed ^ 0x5bd1e995u));
      c->offy[k] = (float)(b0 + (b1 - b0) * (double)or_rng(0xffffffffu - k, 1u, c->cfg->seed ^ 0x5bd1e995u));
    }
}

/* PerspectiveCamera::generateRay (camera.cpp:42-56), plus a shutter time
 * that moves the camera from pos to pos1 (motion blur; zero shutter is the
 * reference's static camera). */
static void camera_ray(const rctx *c, uint32_t i, uint32_t j, uint32_t k, uint32_t g, v3 *o,
                       v3 *d) {
  const or_scene *s = c->s;
  const pt_camera_desc *cam = &s->cam;
  float W = (float)c->cfg->width, H = (float)c->cfg->height;
  float u = ((float)i + c->offx[k]) / W;
  float v = ((float)j + c->offy[k]) / H;
  float radius = (cam->aperture > 0.f && isfinite(cam->aperture)) ? cam->focal_dist / cam->aperture * 0.5f : 0.f;
  float lu = -radius + (radius - -radius) * or_rng(g, slot(0, SLOT_LENS_X), c->cfg->seed);
  float lv = -radius + (radius - -radius) * or_rng(g, slot(0, SLOT_LENS_Y), c->cfg->seed);
  v3 pos = V3(cam->pos[0], cam->pos[1], cam->pos[2]);
  if (cam->shutter > 0.f) {
    float t = or_rng(g, slot(0, SLOT_TIME), c->cfg->seed) * cam->shutter;
    v3 p1 = V3(cam->pos1[0], cam->pos1[1], cam->pos1[2]);
    pos = vadd(pos, vscale(vsub(p1, pos), t));
  }
  v3 front = V3(cam->front[0], cam->front[1], cam->front[2]);
  v3 lookat = vneg(vnorm(front));
  v3 bl = vsub(vsub(pos, vdivs(s->cam_horiz, 2.0f)), vdivs(s->cam_vert, 2.0f));
  bl = vsub(bl, vscale(lookat, cam->focal_dist));
  v3 ps = vadd(pos, vadd(vscale(s->cam_h, lu), vscale(s->cam_v, lv)));
  v3 dir = vsub(vadd(vadd(bl, vscale(s->cam_horiz, u)), vscale(s->cam_vert, v)), ps);
  if (cam->normalize_dir) dir = vnorm(dir);
  *o = ps;
  *d = dir;
}

int or_camera_rays(const or_scene *s, const pt_render_cfg *cfg, const uint32_t *pixel,
                   const uint32_t *k, size_t n, float *rays) {
  rctx c;
  memset(&c, 0, sizeof(c));
  c.s = s;
  c.cfg = cfg;
  c.spp = cfg->spp_x * cfg->spp_y;
  make_offsets(&c);
  for (size_t q = 0; q < n; ++q) {
    uint32_t i = pixel[q] % cfg->width, j = pixel[q] / cfg->width;
    uint32_t g = 1u + k[q] + c.spp * (j + cfg->height * i);
    v3 o, d;
    camera_ray(&c, i, j, k[q], g, &o, &d);
    float *r = &rays[8 * q];
    r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = d.x; r[4] = d.y; r[5] = d.z;
    r[6] = 1e-4f; r[7] = 1e5f;
  }
  free(c.offx);
  free(c.offy);
  return 0;
}

static v3 miss_radiance(const or_scene *s, const pt_render_cfg *cfg, v3 d) {
  if (s->env) return or_env_lookup(s, vnorm(d));
  return V3(cfg->bg[0], cfg->bg[1], cfg->bg[2]);
}

/* The NEE loop of getRayColor (maxdepth_nee.hpp:90-232). */
static v3 nee_reference(const rctx *c, const or_hit *h, v3 wo, const pt_material_desc *m,
                        uint32_t g, uint32_t d, uint64_t *cnt) {
  const or_scene *s = c->s;
  v3 lt = vs(0.f);
  int N = s->n_lights;
  if (N <= 0 || c->ns == 0) return lt;
  for (uint32_t si = 0; si < c->ns; ++si) {
    uint32_t ii = si / c->ly, jj = si % c->ly; /* getOffsets(LX, LY) (random.hpp:47-63) */
    float a0 = (float)ii / (float)c->lx, a1 = (float)(ii + 1) / (float)c->lx;
    float b0 = (float)jj / (float)c->ly, b1 = (float)(jj + 1) / (float)c->ly;
    float ox = (float)(a0 + (a1 - a0) * (double)or_rng(g, slot(d, SLOT_NEE_OFF + 2 * si), c->cfg->seed));
    float oy = (float)(b0 + (b1 - b0) * (double)or_rng(g, slot(d, SLOT_NEE_OFF + 2 * si + 1), c->cfg->seed));
    /* LightList::uniform_get (lightList.hpp:31-35) */
    int li = (int)(or_rng(g, slot(d, SLOT_NEE_LIGHT + si), c->cfg->seed) * (float)N) % N;
    const or_light *L = &s->lights[li];
    v3 wi, le;
    float pdf, fac;
    int vis;
    if (L->type == OR_LIGHT_ENV) {
      v3 nl;
      v3 lp = or_light_point(s, L, ox, oy, &nl);
      wi = vnorm(vsub(lp, h->p));
      if (cnt) cnt[2]++;
      vis = !or_scene_intersect(s, h->p, wi, 0.001f, FLT_MAX, 1, NULL, NULL);
      if (vis) {
        le = or_env_lookup(s, wi);
        pdf = (1.0f / (float)N) * (float)(1.f / (4.f * M_PI));
        fac = fmaxf(0.f, vdot(h->n, wi));
      } else {
        le = vs(0.f); pdf = 1.f; fac = 0.f;
      }
    } else {
      v3 nl;
      v3 lp = or_light_point(s, L, ox, oy, &nl);
      v3 ld = vsub(lp, h->p);
      float len = vlen(ld);
      ld = vdivs(ld, len);
      wi = ld;
      if (len < 0.001f) {
        vis = 0;
      } else {
        if (cnt) cnt[2]++;
        vis = !or_scene_intersect(s, h->p, ld, 0.001f, len - 0.001f, 1, NULL, NULL);
      }
      if (vis) {
        le = (vdot(nl, vneg(wi)) > 0.f) ? L->le : vs(0.f);
        pdf = (L->area <= 1e-7f) ? 0.f : 1.f / (L->area * (float)N);
        float noi = fmaxf(0.f, vdot(h->n, wi));
        float nol = fmaxf(0.f, vdot(nl, vneg(wi)));
        float d2 = len * len;
        fac = (d2 < 1e-7f) ? 0.f : (nol * noi) / d2;
      } else {
        le = vs(0.f); pdf = 1.f; fac = 0.f;
      }
    }
    if (vis && pdf > 1e-7f) {
      v3 f = or_bsdf_f(m, wo, wi, h);
      lt = vadd(lt, vdivs(vscale(vmul(f, le), fac), pdf));
    }
  }
  return vdivs(lt, (float)c->ns);
}

/* Solid-angle pdf with which the NEE+MIS estimator samples point p on
 * light li from o. */
static float light_pdf_sa(const or_scene *s, int li, v3 o, v3 p) {
  const or_light *L = &s->lights[li];
  if (L->type == OR_LIGHT_ENV) return or_env_pdf_dir(s, vsub(p, o));
  v3 w = vsub(p, o);
  float d2 = vdot(w, w);
  v3 nl = (L->type == OR_LIGHT_QUAD) ? L->n : vnorm(vsub(p, L->c));
  float c = fabsf(vdot(nl, vnorm(w)));
  if (!(c > 0.f) || !(L->area > 0.f)) return 0.f;
  return d2 / (L->area * c);
}

static inline float power2(float a, float b) { return (a * a) / (a * a + b * b); }

/* One path sample of the mixture-pdf integrators.  getRayColor of
 * maxdepth_naive.hpp:9-75 (roulette = 0) and roulette_naive.hpp:9-78
 * (roulette = 1), unrolled:
 *   - a miss returns bg_color (the environment light is not consulted);
 *   - the depth limit returns bg_color (naive) or 0 (roulette);
 *   - Material::emit is added at every hit (DiffuseEmitter: two-sided);
 *   - roulette: q = clamp(maxComponent(attenuation), 0.5, 1) right after a
 *     successful scatter, the path ends when rand > q, else f /= q;
 *   - specular lobes: throughput *= attenuation;
 *   - otherwise wi is drawn from MixedPDF (mixedPDF.hpp) of light_samples
 *     HittablePDFs towards uniformly chosen area lights (the
 *     ObjectList(LightList) shapes, objectList.hpp:15-60) plus the
 *     material's pdf, equal weights: throughput *= f cos / pdf_mix.
 * The environment light is not among the light objects (it is appended to
 * the LightList after them, renderManager.hpp:88-95).  Without area lights
 * the mixture is the material's pdf alone (the reference would index an
 * empty list). */
static v3 trace_path_naive(const rctx *c, uint32_t i, uint32_t j, uint32_t k, int roulette, uint64_t *cnt) {
  const or_scene *s = c->s;
  const pt_render_cfg *cfg = c->cfg;
  uint32_t g = 1u + k + c->spp * (j + cfg->height * i);
  uint32_t pixel = j * cfg->width + i;
  v3 ro, rd;
  camera_ray(c, i, j, k, g, &ro, &rd);
  v3 L = vs(0.f), beta = vs(1.f);
  const v3 bg = V3(cfg->bg[0], cfg->bg[1], cfg->bg[2]);
  const uint32_t D = cfg->max_depth;
  const int NA = s->n_lights - (s->env ? 1 : 0);  /* area lights come first */
  const uint32_t M = cfg->light_samples < 4u ? cfg->light_samples : 4u;
  if (cnt) cnt[0]++;
  for (uint32_t d = 0;; ++d) {
    if (d >= D) {
      if (!roulette) L = vadd(L, vmul(beta, bg));
      break;
    }
    or_hit h;
    if (cnt) cnt[1]++;
    if (!or_scene_intersect(s, ro, rd, 1e-4f, 1e5f, 0, &h, cnt ? cnt + 4 : NULL)) {
      L = vadd(L, vmul(beta, bg));
      break;
    }
    const pt_material_desc *m = &s->mats[h.mat];
    float uc = or_sample_dim(cfg->sampler, cfg->seed, cfg->spp_x, cfg->spp_y, pixel, k, g, 3 * d);
    float u0 = or_sample_dim(cfg->sampler, cfg->seed, cfg->spp_x, cfg->spp_y, pixel, k, g, 3 * d + 1);
    float u1 = or_sample_dim(cfg->sampler, cfg->seed, cfg->spp_x, cfg->spp_y, pixel, k, g, 3 * d + 2);
    float aux = or_rng(g, slot(d, SLOT_AUX), cfg->seed);
    or_srec sr;
    if (!or_scatter(m, rd, &h, uc, u0, u1, aux, &sr)) {
      L = vadd(L, vmul(beta, or_emit(m, rd, &h)));
      break;
    }
    float q = 1.f;
    if (roulette) {
      q = fminf(fmaxf(vmaxc(sr.att), 0.5f), 1.f);
      if (or_rng(g, slot(d, SLOT_RR), cfg->seed) > q) break;
    }
    L = vadd(L, vmul(beta, or_emit(m, rd, &h)));
    if (sr.flags & OR_SPECULAR) {
      beta = vmul(beta, vdivs(sr.att, q));
      ro = h.p;
      rd = sr.wi;
      continue;
    }
    v3 wo = vneg(vnorm(rd));
    /* the light objects of the mixture, then one component */
    int li[4] = {0, 0, 0, 0};
    const uint32_t nl = NA > 0 ? M : 0u;
    for (uint32_t si = 0; si < nl; ++si)
      li[si] = (int)(or_rng(g, slot(d, SLOT_NEE_LIGHT + si), cfg->seed) * (float)NA) % NA;
    const uint32_t ncomp = nl + 1u;
    uint32_t comp = (uint32_t)floorf(or_rng(g, slot(d, SLOT_MIX), cfg->seed) * (float)ncomp) % ncomp;
    v3 wi;
    if (comp < nl) {
      float e0 = or_rng(g, slot(d, SLOT_NEE_OFF), cfg->seed);
      float e1 = or_rng(g, slot(d, SLOT_NEE_OFF + 1), cfg->seed);
      wi = vnorm(or_shape_sample(s, &s->lights[li[comp]], h.p, e0, e1));
    } else {
      wi = vnorm(sr.wi);
    }
    float pdf = 0.f;
    const float wgt = 1.f / (float)ncomp;
    for (uint32_t si = 0; si < nl; ++si) pdf += wgt * or_shape_pdf(s, &s->lights[li[si]], h.p, wi);
    pdf += wgt * or_bsdf_pdf(m, wo, wi, &h);
    v3 f = vdivs(or_bsdf_f(m, wo, wi, &h), q);
    float cs = fabsf(vdot(h.n, wi));
    beta = vscale(vmul(beta, f), cs / pdf);
    ro = h.p;
    rd = wi;
    if (!(beta.x != 0.f || beta.y != 0.f || beta.z != 0.f)) break;  /* nothing more can arrive */
  }
  return L;
}

/* One path sample of getRayColor in integrator/maxdepth_reservoir_di.hpp:9-131,
 * unrolled:
 *   - max_depth == 0 and misses return bg_color;
 *   - after a successful scatter the continuation direction comes from the
 *     material's pdf, or from CosinePDF(n) when it has none (specular
 *     lobes), with weight attenuation * cos / pdf (:39-50, 126-129);
 *   - direct light is resampled from light_samples candidates (light,
 *     jittered point): contribution f Le A G V with the visibility inside,
 *     reservoir weight |contribution| (vector length), streaming selection
 *     rand < w / W, estimate sample * W / light_samples (:52-100);
 *   - when the continuation ray hits an emitter the vertex returns 0: its
 *     direct light and everything after it are dropped (:113-115);
 *   - the environment light is a candidate like the others, with area
 *     +infinity and its texture read at the candidate's offset (so such a
 *     sample is not finite, and the film's NaN rule drops it, as the
 *     reference's pixel sum does). */
static v3 trace_path_reservoir(const rctx *c, uint32_t i, uint32_t j, uint32_t k, uint64_t *cnt) {
  const or_scene *s = c->s;
  const pt_render_cfg *cfg = c->cfg;
  uint32_t g = 1u + k + c->spp * (j + cfg->height * i);
  uint32_t pixel = j * cfg->width + i;
  v3 ro, rd;
  camera_ray(c, i, j, k, g, &ro, &rd);
  v3 L = vs(0.f), beta = vs(1.f);
  const v3 bg = V3(cfg->bg[0], cfg->bg[1], cfg->bg[2]);
  const uint32_t D = cfg->max_depth;
  const int N = s->n_lights;
  const uint32_t ns = c->ns < 4u ? c->ns : 4u;
  v3 pending = vs(0.f);  /* the previous vertex's direct light, times its throughput */
  if (cnt) cnt[0]++;
  for (uint32_t d = 0;; ++d) {
    or_hit h;
    if (cnt) cnt[1]++;
    int hit = or_scene_intersect(s, ro, rd, 1e-4f, 1e5f, 0, &h, cnt ? cnt + 4 : NULL);
    if (d > 0 && hit && or_is_emitter(&s->mats[h.mat])) break;  /* :113-115 */
    L = vadd(L, pending);
    pending = vs(0.f);
    if (d >= D || !hit) {
      L = vadd(L, vmul(beta, bg));
      break;
    }
    const pt_material_desc *m = &s->mats[h.mat];
    float uc = or_sample_dim(cfg->sampler, cfg->seed, cfg->spp_x, cfg->spp_y, pixel, k, g, 3 * d);
    float u0 = or_sample_dim(cfg->sampler, cfg->seed, cfg->spp_x, cfg->spp_y, pixel, k, g, 3 * d + 1);
    float u1 = or_sample_dim(cfg->sampler, cfg->seed, cfg->spp_x, cfg->spp_y, pixel, k, g, 3 * d + 2);
    float aux = or_rng(g, slot(d, SLOT_AUX), cfg->seed);
    or_srec sr;
    if (!or_scatter(m, rd, &h, uc, u0, u1, aux, &sr)) {
      L = vadd(L, vmul(beta, or_emit(m, rd, &h)));
      break;
    }
    v3 wo = vneg(vnorm(rd));
    v3 wi;
    float pdf;
    if (sr.flags & OR_SPECULAR) { /* CosinePDF(n): get(uc, u) and at(dir) */
      wi = or_cosine_dir(h.n, u0, u1);
      pdf = or_cosine_pdf(h.n, wi);
    } else {
      wi = vnorm(sr.wi);
      pdf = or_bsdf_pdf(m, wo, wi, &h);
    }
    float cs = fabsf(vdot(h.n, wi));
    /* resampled direct light */
    v3 sample = vs(0.f);
    float W = 0.f;
    if (N > 0) {
      for (uint32_t si = 0; si < ns; ++si) {
        uint32_t ii = si / c->ly, jj = si % c->ly;
        float a0 = (float)ii / (float)c->lx, a1 = (float)(ii + 1) / (float)c->lx;
        float b0 = (float)jj / (float)c->ly, b1 = (float)(jj + 1) / (float)c->ly;
        float ox = a0 + (a1 - a0) * or_rng(g, slot(d, SLOT_NEE_OFF + 2 * si), cfg->seed);
        float oy = b0 + (b1 - b0) * or_rng(g, slot(d, SLOT_NEE_OFF + 2 * si + 1), cfg->seed);
        int li = (int)(or_rng(g, slot(d, SLOT_NEE_LIGHT + si), cfg->seed) * (float)N) % N;
        const or_light *Lg = &s->lights[li];
        v3 nl;
        v3 lp = or_light_point(s, Lg, ox, oy, &nl);
        v3 ld = vsub(lp, h.p);
        v3 lwi = vnorm(ld);
        float len = vlen(ld);
        if (cnt) cnt[2]++;
        int occ = or_scene_intersect(s, h.p, lwi, 0.001f, len - 0.001f, 1, NULL, NULL);
        float noi = fmaxf(vdot(h.n, lwi), 0.f);
        float nol = fmaxf(vdot(nl, vneg(lwi)), 0.f);
        float G = nol * noi / vdot(ld, ld);
        float V = occ ? 0.f : 1.f;
        v3 f = or_bsdf_f(m, wo, lwi, &h);
        v3 le = Lg->type == OR_LIGHT_ENV ? or_env_lookup_uv(s, ox, oy) : Lg->le;
        float area = Lg->type == OR_LIGHT_ENV ? INFINITY : Lg->area;
        v3 cand = vscale(vmul(f, le), area * G * V);
        float w = vlen(cand);
        W += w;
        if (or_rng(g, slot(d, SLOT_RIS + si), cfg->seed) < w / W) sample = cand;
      }
    }
    pending = vmul(beta, vscale(sample, W / (float)(ns ? ns : 1)));
    beta = vscale(vmul(beta, sr.att), cs / pdf);
    ro = h.p;
    rd = wi;
  }
  return L;
}

/* One path sample.  Returns the radiance. */
/* HenyeyGreenstein (medium/phaseMath.hpp:7-15): (1 - g^2) / (4 pi (1 + g^2 +
 * 2 g cos)^1.5), the PBRT (wo, wi) sign convention; INV4PI is 1 / (4 M_PI)
 * rounded to float. */
static float hg_phase(float cos_theta, float g) {
  float upper = 1.f - g * g;
  float lower = 1.f + g * g + 2.f * g * cos_theta;
  lower *= sqrtf(fmaxf(0.f, lower));
  float inv4pi = (float)(1.0 / (4.0 * M_PI));
  return upper * inv4pi / lower;
}

/* One path sample of getRayColor in integrator/maxdepth_volume.hpp:72-118,
 * single scattering in the scene's global HomogeneousMedium, with
 * estimate_l_scatter1 (:10-70):
 *   - no medium: 0;
 *   - the camera ray's closest hit (HitRecord::t = FLT_MAX on a miss);
 *   - a free-flight distance t = -log(1 - u) / max(sigma_t), u the sample's
 *     first dimension (halton_sampler.get1D());
 *   - t before the hit: (sigma_s / sigma_t) * L_s1 / max(sigma_t), where L_s1
 *     is the mean over LIGHT_SAMPLE_NUM light samples: a uniformly chosen
 *     light, a point p' on it, the shadow ray from p' to the scatter point
 *     (0.001, dist - 0.001), Le = L_emit towards the scatter point (one-sided),
 *     |n' . dir| * HG(-d . -dir) * exp(-sigma_t dist) / pdf / dist^2 with pdf =
 *     1 / (area N); samples closer than gl::epsilon = 1e-3 are skipped;
 *   - otherwise the surface's emission (Material::emit), else 0.
 * The reference divides by max(sigma_t) once more than the free-flight pdf
 * asks for; that is kept. Light choice and points use the counter stream of
 * vertex 0 (D-1). */
static v3 trace_path_volume(const rctx *c, uint32_t i, uint32_t j, uint32_t k, uint64_t *cnt) {
  const or_scene *s = c->s;
  const pt_render_cfg *cfg = c->cfg;
  uint32_t g = 1u + k + c->spp * (j + cfg->height * i);
  uint32_t pixel = j * cfg->width + i;
  v3 ro, rd;
  camera_ray(c, i, j, k, g, &ro, &rd);
  if (cnt) cnt[0]++;
  if (!s->has_medium) return vs(0.f);
  or_hit h;
  if (cnt) cnt[1]++;
  int hit = or_scene_intersect(s, ro, rd, 1e-4f, 1e5f, 0, &h, cnt ? cnt + 4 : NULL);
  float th = hit ? h.t : FLT_MAX;
  v3 st = vadd(s->med_sa, s->med_ss);
  float smax = vmaxc(st);
  float u = or_sample_dim(cfg->sampler, cfg->seed, cfg->spp_x, cfg->spp_y, pixel, k, g, 0);
  float tf = -logf(1.f - u) / smax;
  if (tf < th) {
    v3 ps = vadd(ro, vscale(rd, tf));
    v3 wo = vneg(vnorm(rd));
    const int N = s->n_lights;
    const uint32_t ns = cfg->light_samples < 4u ? cfg->light_samples : 4u;
    v3 alb = V3(s->med_ss.x / st.x, s->med_ss.y / st.y, s->med_ss.z / st.z);
    v3 acc = vs(0.f);
    for (uint32_t si = 0; N > 0 && si < ns; ++si) {
      int li = (int)(or_rng(g, slot(0, SLOT_NEE_LIGHT + si), cfg->seed) * (float)N) % N;
      const or_light *Lg = &s->lights[li];
      v3 nl;
      v3 lp = or_light_point(s, Lg, or_rng(g, slot(0, SLOT_NEE_OFF + 2 * si), cfg->seed),
                             or_rng(g, slot(0, SLOT_NEE_OFF + 2 * si + 1), cfg->seed), &nl);
      v3 dv = vsub(ps, lp);
      v3 dl = vnorm(dv);
      float dist = vlen(dv);
      if (cnt) cnt[2]++;
      int vis = !or_scene_intersect(s, lp, dl, 0.001f, dist - 0.001f, 1, NULL, NULL);
      if (!vis || dist < 1e-3f) continue;
      v3 le = (vdot(nl, dl) > 0.f) ? Lg->le : vs(0.f);
      float pdf = 1.f / (Lg->area * (float)N);
      float cosl = fabsf(vdot(nl, dl));
      float ph = hg_phase(vdot(wo, vneg(dl)), s->med_g);
      v3 T = V3(expf(-st.x * dist), expf(-st.y * dist), expf(-st.z * dist));
      v3 cc = vdivs(vdivs(vmul(vscale(le, cosl * ph), T), pdf), dist * dist);
      acc = vadd(acc, cc);
    }
    if (ns > 0) acc = vdivs(acc, (float)ns);
    return vdivs(vmul(alb, acc), smax);
  }
  if (hit) {
    const pt_material_desc *m = &s->mats[h.mat];
    if (m->kind == PT_MAT_EMITTER) return or_emit(m, rd, &h);
  }
  return vs(0.f);
}

static v3 trace_path(const rctx *c, uint32_t i, uint32_t j, uint32_t k, uint64_t *cnt) {
  if (c->cfg->estimator == PT_EST_VOLUME) return trace_path_volume(c, i, j, k, cnt);
  if (c->cfg->estimator == PT_EST_NAIVE || c->cfg->estimator == PT_EST_ROULETTE)
    return trace_path_naive(c, i, j, k, c->cfg->estimator == PT_EST_ROULETTE, cnt);
  if (c->cfg->estimator == PT_EST_RESERVOIR) return trace_path_reservoir(c, i, j, k, cnt);
  const or_scene *s = c->s;
  const pt_render_cfg *cfg = c->cfg;
  uint32_t g = 1u + k + c->spp * (j + cfg->height * i);
  uint32_t pixel = j * cfg->width + i;
  v3 ro, rd;
  camera_ray(c, i, j, k, g, &ro, &rd);
  v3 L = vs(0.f), beta = vs(1.f);
  or_hit h;
  int have = 0, hit = 0;
  const uint32_t D = cfg->max_depth;
  const int guided = cfg->estimator == OR_EST_GUIDED;
  const int mis = cfg->estimator == PT_EST_NEE_MIS || guided;
  int prev_spec = 1;
  float prev_pdf = 0.f;
  v3 prev_p = ro;
  const int N = s->n_lights;
  if (cnt) cnt[0]++;
  for (uint32_t d = 0;; ++d) {
    if (!mis && d >= D) break; /* max_depth == 0 returns 0 (maxdepth_nee.hpp:16-17) */
    if (!have) {
      if (cnt) cnt[1]++;
      hit = or_scene_intersect(s, ro, rd, 1e-4f, 1e5f, 0, &h, cnt ? cnt + 4 : NULL);
    }
    have = 0;
    if (!hit) {
      if (!mis) {
        L = vadd(L, vmul(beta, miss_radiance(s, cfg, rd)));
      } else if (prev_spec) {
        if (d < D || d == 0) L = vadd(L, vmul(beta, miss_radiance(s, cfg, rd)));
      } else if (s->env) {
        float pl = or_env_pdf_dir(s, rd) / (float)N;
        L = vadd(L, vscale(vmul(beta, or_env_lookup(s, rd)), power2(prev_pdf, pl)));
      } else {
        L = vadd(L, vmul(beta, miss_radiance(s, cfg, rd)));
      }
      break;
    }
    const pt_material_desc *m = &s->mats[h.mat];
    if (mis && or_is_emitter(m)) {
      int li = or_prim_light(s, h.prim);
      if (prev_spec || li < 0) {
        if (d < D) L = vadd(L, vmul(beta, or_emit(m, rd, &h)));
      } else {
        const or_light *Lg = &s->lights[li];
        v3 le = (vdot(h.ng_light, vneg(vnorm(rd))) > 0.f) ? Lg->le : vs(0.f);
        float pl = light_pdf_sa(s, li, prev_p, h.p) / (float)N;
        L = vadd(L, vscale(vmul(beta, le), power2(prev_pdf, pl)));
      }
      break;
    }
    if (mis && d >= D) break;
    float uc = or_sample_dim(cfg->sampler, cfg->seed, cfg->spp_x, cfg->spp_y, pixel, k, g, 3 * d);
    float u0 = or_sample_dim(cfg->sampler, cfg->seed, cfg->spp_x, cfg->spp_y, pixel, k, g, 3 * d + 1);
    float u1 = or_sample_dim(cfg->sampler, cfg->seed, cfg->spp_x, cfg->spp_y, pixel, k, g, 3 * d + 2);
    float aux = or_rng(g, slot(d, SLOT_AUX), cfg->seed);
    or_srec sr;
    /* guided: the technique is drawn first, the mixture with probability c */
    const float gc = (guided && d < s->gverts && or_guidable(m, s->gkinds)) ? s->gmix[40] : 0.f;
    if (gc > 0.f && or_rng(g, slot(d, SLOT_GUIDE), cfg->seed) < gc) {
      sr.wi = or_guide_sample(s->gmix, or_rng(g, slot(d, SLOT_GUIDE + 1), cfg->seed),
                              or_rng(g, slot(d, SLOT_GUIDE + 2), cfg->seed),
                              or_rng(g, slot(d, SLOT_GUIDE + 3), cfg->seed));
      sr.flags = 0;
      sr.att = vs(0.f);
      sr.pdf = 0.f;
    } else if (!or_scatter(m, rd, &h, uc, u0, u1, aux, &sr)) {
      L = vadd(L, vmul(beta, or_emit(m, rd, &h)));
      break;
    }
    if (sr.flags & OR_SPECULAR) {
      /* maxdepth_nee.hpp:73-78 (and the split-ray case taken stochastically) */
      beta = vmul(beta, sr.att);
      ro = h.p;
      rd = sr.wi;
      prev_spec = 1;
    } else {
      v3 wo = vneg(vnorm(rd));
      v3 wi = vnorm(sr.wi);
      float cs = fabsf(vdot(h.n, wi));
      float pdf = sr.pdf;
      const v3 vp = h.p; /* this vertex (h is overwritten by the continuation's hit below) */
      if (!mis) {
        L = vadd(L, vmul(beta, nee_reference(c, &h, wo, m, g, d, cnt)));
        if (cnt) cnt[1]++;
        hit = or_scene_intersect(s, h.p, wi, 1e-4f, 1e5f, 0, &h, cnt ? cnt + 4 : NULL);
        if (hit && or_is_emitter(&s->mats[h.mat])) break; /* maxdepth_nee.hpp:246-247 */
        /* the recursion traverses the same ray again (maxdepth_nee.hpp:252-255) */
        have = !(cfg->flags & OR_RENDER_REFERENCE_TRAVERSALS);
      } else if (N > 0) {
        int li = (int)(or_rng(g, slot(d, SLOT_MIS_LIGHT), cfg->seed) * (float)N) % N;
        const or_light *Lg = &s->lights[li];
        float e0 = or_rng(g, slot(d, SLOT_NEE_OFF), cfg->seed);
        float e1 = or_rng(g, slot(d, SLOT_NEE_OFF + 1), cfg->seed);
        v3 lwi, le = vs(0.f);
        float pl = 0.f;
        int vis = 0;
        if (Lg->type == OR_LIGHT_ENV) {
          float euv[2];
          lwi = or_env_sample_dir(s, e0, e1, &pl, euv);
          if (pl > 0.f) {
            if (cnt) cnt[2]++;
            vis = !or_scene_intersect(s, h.p, lwi, 0.001f, FLT_MAX, 1, NULL, NULL);
            /* the radiance at the sampled texture coordinates (the direction's own) */
            le = euv[0] >= 0.f ? or_env_lookup_uv(s, euv[0], euv[1]) : or_env_lookup(s, lwi);
          }
        } else {
          v3 nl;
          v3 lp = or_light_point(s, Lg, e0, e1, &nl);
          v3 ld = vsub(lp, h.p);
          float len = vlen(ld);
          lwi = vdivs(ld, len);
          float nol = vdot(nl, vneg(lwi));
          if (len >= 0.001f && nol > 0.f && Lg->area > 0.f) {
            pl = len * len / (Lg->area * nol);
            if (cnt) cnt[2]++;
            vis = !or_scene_intersect(s, h.p, lwi, 0.001f, len - 0.001f, 1, NULL, NULL);
            le = Lg->le;
          }
        }
        // a reflection-only BSDF takes no light from behind its surface (the
        // NEE integrator's max(0, n.wi), maxdepth_nee.hpp:137, 192)
        if (reflect_only(m->kind) && vdot(h.n, lwi) <= 0.f) pl = 0.f;
        if (vis && pl > 0.f) {
          pl /= (float)N;
          v3 f = or_bsdf_f(m, wo, lwi, &h);
          float pb = or_bsdf_pdf(m, wo, lwi, &h);
          if (gc > 0.f) pb = gc * or_guide_pdf(s->gmix, lwi) + (1.f - gc) * pb;
          float w = power2(pl, pb);
          float cl = fabsf(vdot(h.n, lwi));
          L = vadd(L, vscale(vmul(beta, vmul(f, le)), cl / pl * w));
        }
      }
      if (gc > 0.f) {
        /* one-sample MIS over the two techniques; a direction behind a
         * reflection-only surface carries nothing */
        const int below = reflect_only(m->kind) && vdot(h.n, wi) <= 0.f;
        const v3 f = below ? vs(0.f) : or_bsdf_f(m, wo, wi, &h);
        pdf = gc * or_guide_pdf(s->gmix, wi) + (1.f - gc) * or_bsdf_pdf(m, wo, wi, &h);
        beta = pdf > 0.f ? vscale(vmul(beta, f), cs / pdf) : vs(0.f);
        if (beta.x == 0.f && beta.y == 0.f && beta.z == 0.f) break;
      } else {
        beta = vscale(vmul(beta, sr.att), cs / pdf);
      }
      prev_pdf = pdf;
      prev_p = vp;
      prev_spec = 0;
      ro = vp;
      rd = wi;
    }
    if (cfg->rr_depth && d >= cfg->rr_depth) {
      float q = fminf(1.f, vmaxc(beta));
      if (!(or_rng(g, slot(d, SLOT_RR), cfg->seed) < q)) break;
      beta = vdivs(beta, q);
    }
  }
  return L;
}

static int setup(rctx *c, const or_scene *s, const pt_render_cfg *cfg) {
  memset(c, 0, sizeof(*c));
  if (!s->committed || !s->has_cam) return -1;
  c->s = s;
  c->cfg = cfg;
  c->spp = cfg->spp_x * cfg->spp_y;
  c->ns = cfg->light_samples;
  grid_factor(c->ns ? c->ns : 1, &c->lx, &c->ly);
  make_offsets(c);
  return 0;
}

int or_render(const or_scene *s, const pt_render_cfg *cfg, float *out_rgb, float *out_var,
              float *out_samples, uint64_t *counters, int threads) {
  rctx c;
  if (setup(&c, s, cfg)) return -1;
  uint32_t W = cfg->width, H = cfg->height;
  uint32_t k0 = cfg->spp_begin, kn = cfg->spp_count ? cfg->spp_count : c.spp - k0;
  uint64_t tot[8] = {0};
  int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
  {
    uint64_t cnt[8] = {0};
#pragma omp for schedule(dynamic, 16)
    for (long long p = 0; p < (long long)W * H; ++p) {
      uint32_t i = (uint32_t)(p % W), j = (uint32_t)(p / W);
      double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
      for (uint32_t k = k0; k < k0 + kn; ++k) {
        v3 Lp = trace_path(&c, i, j, k, counters ? cnt : NULL);
        if (!v3_isfinite(Lp)) {
          cnt[3]++;
          Lp = vs(0.f);
        }
        if (out_samples) {
          float *o = &out_samples[3 * ((size_t)p * kn + (k - k0))];
          o[0] = Lp.x; o[1] = Lp.y; o[2] = Lp.z;
        }
        sum[0] += Lp.x; sum[1] += Lp.y; sum[2] += Lp.z;
        sq[0] += (double)Lp.x * Lp.x; sq[1] += (double)Lp.y * Lp.y; sq[2] += (double)Lp.z * Lp.z;
      }
      for (int ch = 0; ch < 3; ++ch) {
        double mean = sum[ch] / kn;
        out_rgb[3 * p + ch] = (float)mean;
        if (out_var) out_var[3 * p + ch] = kn > 1 ? (float)((sq[ch] - kn * mean * mean) / (kn - 1)) : 0.f;
      }
    }
#pragma omp critical(or_render_cnt)
    for (int q = 0; q < 8; ++q) tot[q] += cnt[q];
  }
  if (counters) memcpy(counters, tot, sizeof(tot));
  free(c.offx);
  free(c.offy);
  return 0;
}

double or_render_range(const or_scene *s, const pt_render_cfg *cfg, uint64_t first,
                       uint64_t count, int threads, uint64_t *counters) {
  rctx c;
  if (setup(&c, s, cfg)) return -1.0;
  uint32_t W = cfg->width;
  double lum = 0.0;
  uint64_t tot[8] = {0};
  int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt) reduction(+ : lum)
  {
    uint64_t cnt[8] = {0};
#pragma omp for schedule(dynamic, 64)
    for (long long q = 0; q < (long long)count; ++q) {
      uint64_t sidx = first + (uint64_t)q;
      uint64_t p = sidx / c.spp;
      uint32_t k = (uint32_t)(sidx % c.spp);
      uint32_t i = (uint32_t)(p % W), j = (uint32_t)(p / W);
      v3 Lp = trace_path(&c, i, j, k, counters ? cnt : NULL);
      if (v3_isfinite(Lp)) lum += (double)(Lp.x + Lp.y + Lp.z);
    }
#pragma omp critical(or_range_cnt)
    for (int z = 0; z < 8; ++z) tot[z] += cnt[z];
  }
  if (counters) memcpy(counters, tot, sizeof(tot));
  free(c.offx);
  free(c.offy);
  return lum;
}
