// pt_kernels.hip -- wavefront path-tracing kernels for gfx950.
//
// One render is a sequence of chunks of path samples (pixel-major order,
// whole pixels per chunk).  Per chunk:
//   k_generate                       camera rays for every slot
//   per vertex d:  k_extend          closest hit of the active-ray queue
//                  k_shade<EST>      material, emission, NEE shadow rays,
//                                    continuation rays (queue compaction)
//                  k_shadow          any-hit of the shadow queue, adds the
//                                    unoccluded contributions
//   k_film                           per-pixel sums (fixed order)
// The estimator is getRayColor of integrator/maxdepth_nee.hpp:9-256
// unrolled into a loop (EST 0), or NEE+MIS (EST 1).
#include "pt_device.h"
#include "pt_launch.h"

namespace pt {

__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// ------------------------------------------------------------------ generate
__global__ void __launch_bounds__(256) k_generate(DevScene s, DevCamera cam, RenderParams rp, Wave w,
                                                  const float *offx, const float *offy) {
  uint32_t slot = gtid();
  if (slot >= rp.chunk_paths) return;
  uint64_t q = rp.chunk_first + slot;
  uint32_t pixel = (uint32_t)(q / rp.spp_count);
  uint32_t k = rp.spp_begin + (uint32_t)(q % rp.spp_count);
  uint32_t pi = pixel % rp.width, pj = pixel / rp.width;
  uint32_t g = 1u + k + rp.spp * (pj + rp.height * pi);
  float u = div_rn((float)pi + offx[k], (float)rp.width);
  float v = div_rn((float)pj + offy[k], (float)rp.height);
  f3 o, d;
  camera_ray(cam, u, v, pt_rng(g, rslot(0, SLOT_LENS_X), rp.seed), pt_rng(g, rslot(0, SLOT_LENS_Y), rp.seed),
             pt_rng(g, rslot(0, SLOT_TIME), rp.seed), o, d);
  w.ox[slot] = o.x; w.oy[slot] = o.y; w.oz[slot] = o.z;
  w.dx[slot] = d.x; w.dy[slot] = d.y; w.dz[slot] = d.z;
  w.bx[slot] = 1.f; w.by[slot] = 1.f; w.bz[slot] = 1.f;
  w.lx[slot] = 0.f; w.ly[slot] = 0.f; w.lz[slot] = 0.f;
  w.flags[slot] = 1u;
  w.ppdf[slot] = 0.f;
  w.px[slot] = o.x; w.py[slot] = o.y; w.pz[slot] = o.z;
  w.q[0][slot] = slot;
}

__global__ void k_set_counts(uint32_t *counts, uint32_t q0, uint32_t q1, uint32_t sh) {
  if (threadIdx.x == 0) {
    counts[0] = q0;
    counts[1] = q1;
    counts[2] = sh;
  }
}
__global__ void k_clear_u32(uint32_t *p, uint32_t idx) {
  if (threadIdx.x == 0) p[idx] = 0u;
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
template <bool COUNT>
__global__ void __launch_bounds__(PT_BLOCK) __attribute__((amdgpu_waves_per_eu(PT_TRAV_WAVES))) k_extend(DevScene s, Wave w, int qi) {
  __shared__ int lds[PT_TRAV_LDS(PT_K_EXT, true)];
  const uint32_t n = w.counts[qi];
  const uint32_t *queue = w.q[qi];
  uint32_t nn = 0, np = 0, segs = 0;
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
    uint32_t i = i0 + threadIdx.x;
    if (i < n) {
      uint32_t slot = queue[i];
      f3 o = mk3(w.ox[slot], w.oy[slot], w.oz[slot]);
      f3 d = mk3(w.dx[slot], w.dy[slot], w.dz[slot]);
      float t;
      int prim;
      traverse<false, COUNT, PT_K_EXT>(s, lds, o, d, 1e-4f, 1e5f, t, prim, nn, np);
      w.t[slot] = t;
      w.prim[slot] = prim;
      ++segs;
    }
  }
  if (COUNT) {
    wave_add_u64(&w.stats[CNT_NODES], nn);
    wave_add_u64(&w.stats[CNT_PRIMS], np);
    wave_add_u64(&w.stats[CNT_SEGMENTS], segs);
  }
}

// ------------------------------------------------------------------ shadow
template <bool COUNT>
__global__ void __launch_bounds__(PT_BLOCK) __attribute__((amdgpu_waves_per_eu(PT_TRAV_WAVES))) k_shadow(DevScene s, Wave w) {
  __shared__ int lds[PT_TRAV_LDS(PT_K_SHADOW, false)];
  const uint32_t n = w.counts[CNT_SHADOW];
  uint32_t nn = 0, np = 0, rays = 0;
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
    uint32_t i = i0 + threadIdx.x;
    if (i < n) {
      f3 o = mk3(w.sox[i], w.soy[i], w.soz[i]);
      f3 d = mk3(w.sdx[i], w.sdy[i], w.sdz[i]);
      float t;
      int prim;
      bool occ = traverse<true, COUNT, PT_K_SHADOW>(s, lds, o, d, 0.001f, w.stmax[i], t, prim, nn, np);
      ++rays;
      if (!occ) {
        uint32_t slot = w.sslot[i];
        atomicAdd(&w.lx[slot], w.scr[i]);
        atomicAdd(&w.ly[slot], w.scg[i]);
        atomicAdd(&w.lz[slot], w.scb[i]);
      }
    }
  }
  if (COUNT) {
    wave_add_u64(&w.stats[CNT_SNODES], nn);
    wave_add_u64(&w.stats[CNT_SPRIMS], np);
    wave_add_u64(&w.stats[CNT_SHADOW_RAYS], rays);
  }
}

// ------------------------------------------------------------------ shade
// Appends one shadow ray.  Called from divergent code: the compiler turns the
// uniform-address atomic into one atomic per wave (lane-ordered offsets).
PT_D void emit_shadow(const Wave &w, f3 p, f3 d, float tmax, f3 c, uint32_t slot) {
  uint32_t idx = atomicAdd(&w.counts[CNT_SHADOW], 1u);
  if (idx < w.shadow_cap) {
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

// EST 0: getRayColor (maxdepth_nee.hpp:9-256).  EST 1: NEE + MIS.
// Occupancy target of the shading kernel (waves per SIMD): the register
// allocator spills the rarely used material state to reach it.
#ifndef PT_SHADE_WAVES
#define PT_SHADE_WAVES 4
#endif
template <int EST, uint32_t MS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PT_SHADE_WAVES))) k_shade(DevScene s, RenderParams rp, Wave w, int qi, uint32_t d) {
  const uint32_t n = w.counts[qi];
  const uint32_t *qin = w.q[qi];
  uint32_t *qout = w.q[qi ^ 1];
  uint32_t *cout = &w.counts[qi ^ 1];
  const uint32_t D = rp.max_depth;
  const int N = s.n_lights;
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
    uint32_t i = i0 + threadIdx.x;
    bool active = i < n;
    bool cont = false;
    uint32_t slot = 0;
    f3 p = s3(0.f), wi = s3(0.f);
    if (active) {
      slot = qin[i];
      uint64_t q = rp.chunk_first + slot;
      uint32_t pixel = (uint32_t)(q / rp.spp_count);
      uint32_t k = rp.spp_begin + (uint32_t)(q % rp.spp_count);
      uint32_t pi = pixel % rp.width, pj = pixel / rp.width;
      uint32_t g = 1u + k + rp.spp * (pj + rp.height * pi);
      f3 o = mk3(w.ox[slot], w.oy[slot], w.oz[slot]);
      f3 dir = mk3(w.dx[slot], w.dy[slot], w.dz[slot]);
      float t = w.t[slot];
      int prim = w.prim[slot];
      f3 beta = mk3(w.bx[slot], w.by[slot], w.bz[slot]);
      f3 L = mk3(w.lx[slot], w.ly[slot], w.lz[slot]);
      uint32_t fl = w.flags[slot];
      bool prev_spec = fl & 1u;
      bool done = false;
      if (prim < 0) {
        if (EST == 0) {
          L = add3(L, mul3(beta, miss_radiance(s, rp, dir)));
        } else if (prev_spec) {
          if (d < D) L = add3(L, mul3(beta, miss_radiance(s, rp, dir)));
        } else if (s.has_env) {
          float pl = env_pdf(s, dir) / (float)N;
          L = add3(L, scl3(mul3(beta, env_lookup(s, dir)), power_w(w.ppdf[slot], pl)));
        } else {
          L = add3(L, mul3(beta, miss_radiance(s, rp, dir)));
        }
        done = true;
      }
      SurfHit h;
      const pt_material_desc *mat = nullptr;
      if (!done) {
        h = surface(s, prim, o, dir, t);
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
            float pl = light_pdf_sa(s, Lg, mk3(w.px[slot], w.py[slot], w.pz[slot]), h.p) / (float)N;
            L = add3(L, scl3(mul3(beta, le), power_w(w.ppdf[slot], pl)));
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
        if (!bsdf_scatter<MS>(*mat, wo, h.sp, uc, u0, u1, aux, sc)) {
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
              lwi = env_sample(s, e0, e1, pl);
              le = env_lookup(s, lwi);
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
            if (pl > 0.f) {
              pl /= (float)N;
              f3 f = bsdf_f<MS>(*mat, wo, lwi, h.sp);
              float pb = bsdf_pdf<MS>(*mat, wo, lwi, h.sp);
              float cl = fabsf(dot3(h.sp.n, lwi));
              f3 c = scl3(mul3(beta, mul3(f, le)), cl / pl * power_w(pl, pb));
              if ((c.x > 0.f || c.y > 0.f || c.z > 0.f) || !finite3(c)) emit_shadow(w, p, lwi, tmax, c, slot);
            }
          }
          beta = scl3(mul3(beta, sc.att), cs / sc.pdf);
          wi = wn;
          fl = 0u;
          if (EST == 0) {
            cont = d + 1u < D;  // the emitter check of the next hit happens at the next vertex
          } else {
            cont = d + 1u <= D;
            w.ppdf[slot] = sc.pdf;
            w.px[slot] = p.x; w.py[slot] = p.y; w.pz[slot] = p.z;
          }
        }
        if (cont && rp.rr_depth && d >= rp.rr_depth) {
          float qs = fminf(1.f, maxc3(beta));
          if (!(pt_rng(g, rslot(d, SLOT_RR), rp.seed) < qs)) cont = false;
          else beta = divs3(beta, qs);
        }
      }
      w.lx[slot] = L.x; w.ly[slot] = L.y; w.lz[slot] = L.z;
      if (cont) {
        w.bx[slot] = beta.x; w.by[slot] = beta.y; w.bz[slot] = beta.z;
        w.ox[slot] = p.x; w.oy[slot] = p.y; w.oz[slot] = p.z;
        w.dx[slot] = wi.x; w.dy[slot] = wi.y; w.dz[slot] = wi.z;
        w.flags[slot] = fl;
      }
    }
    // converged: append the continuation
    uint32_t idx = wave_append(cout, cont);
    if (cont) qout[idx] = slot;
  }
}

// ------------------------------------------------------------------ film
__global__ void __launch_bounds__(256) k_film(RenderParams rp, Wave w, float *film, float *samples,
                                              unsigned long long *stats) {
  uint32_t pl = gtid();
  uint32_t npix = rp.chunk_paths / rp.spp_count;
  if (pl >= npix) return;
  uint64_t pixel = rp.chunk_first / rp.spp_count + pl;
  float sr = 0.f, sg = 0.f, sb = 0.f;
  uint32_t nan = 0;
  for (uint32_t k = 0; k < rp.spp_count; ++k) {
    uint32_t slot = pl * rp.spp_count + k;
    float r = w.lx[slot], g = w.ly[slot], b = w.lz[slot];
    if (!(isfinite(r) && isfinite(g) && isfinite(b))) {
      r = g = b = 0.f;
      ++nan;
    }
    if (samples) {
      size_t o = 3 * ((size_t)pixel * rp.spp_count + k);
      samples[o] = r; samples[o + 1] = g; samples[o + 2] = b;
    }
    sr += r; sg += g; sb += b;
  }
  film[3 * pixel] += sr;
  film[3 * pixel + 1] += sg;
  film[3 * pixel + 2] += sb;
  if (nan) atomicAdd(&stats[CNT_NAN], (unsigned long long)nan);
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
  if (any)
    hit = traverse<true, false, PT_K_SHADOW>(s, lds, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), r[6], r[7], t, pr, a, b);
  else
    hit = traverse<false, false, PT_K_EXT>(s, lds, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), r[6], r[7], t, pr, a, b);
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
  f3 r = bsdf_f(m, o, w, h);
  f[3 * i] = r.x; f[3 * i + 1] = r.y; f[3 * i + 2] = r.z;
  pdf[i] = bsdf_pdf(m, o, w, h);
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

// ------------------------------------------------------------------ launchers
static inline uint32_t blocks_for(size_t n, uint32_t bs = 256) { return (uint32_t)((n + bs - 1) / bs); }

void launch_generate(hipStream_t st, const DevScene &s, const DevCamera &c, const RenderParams &rp, const Wave &w,
                     const float *offx, const float *offy) {
  hipLaunchKernelGGL(k_generate, dim3(blocks_for(rp.chunk_paths)), dim3(256), 0, st, s, c, rp, w, offx, offy);
}
void launch_set_counts(hipStream_t st, uint32_t *counts, uint32_t q0, uint32_t q1, uint32_t sh) {
  hipLaunchKernelGGL(k_set_counts, dim3(1), dim3(64), 0, st, counts, q0, q1, sh);
}
void launch_clear_count(hipStream_t st, uint32_t *counts, uint32_t idx) {
  hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, st, counts, idx);
}
void launch_extend(hipStream_t st, const DevScene &s, const Wave &w, int qi, uint32_t max_n, uint32_t grid_cap,
                   bool count) {
  uint32_t g = std::min(blocks_for(max_n), grid_cap);
  if (g == 0) g = 1;
  if (count)
    hipLaunchKernelGGL(k_extend<true>, dim3(g), dim3(PT_BLOCK), 0, st, s, w, qi);
  else
    hipLaunchKernelGGL(k_extend<false>, dim3(g), dim3(PT_BLOCK), 0, st, s, w, qi);
}
void launch_shade(hipStream_t st, const DevScene &s, const RenderParams &rp, const Wave &w, int qi, uint32_t d,
                  uint32_t max_n, uint32_t grid_cap) {
  uint32_t g = std::min(blocks_for(max_n), grid_cap);
  if (g == 0) g = 1;
  // the basic material set when the scene uses nothing else (fewer registers)
  const bool basic = (s.mat_kinds & ~kMatBasic) == 0u;
  if (rp.estimator == PT_EST_NEE_MIS) {
    if (basic)
      hipLaunchKernelGGL((k_shade<1, kMatBasic>), dim3(g), dim3(256), 0, st, s, rp, w, qi, d);
    else
      hipLaunchKernelGGL((k_shade<1, kMatAll>), dim3(g), dim3(256), 0, st, s, rp, w, qi, d);
  } else {
    if (basic)
      hipLaunchKernelGGL((k_shade<0, kMatBasic>), dim3(g), dim3(256), 0, st, s, rp, w, qi, d);
    else
      hipLaunchKernelGGL((k_shade<0, kMatAll>), dim3(g), dim3(256), 0, st, s, rp, w, qi, d);
  }
}
void launch_shadow(hipStream_t st, const DevScene &s, const Wave &w, uint32_t max_n, uint32_t grid_cap, bool count) {
  uint32_t g = std::min(blocks_for(max_n), grid_cap);
  if (g == 0) g = 1;
  if (count)
    hipLaunchKernelGGL(k_shadow<true>, dim3(g), dim3(PT_BLOCK), 0, st, s, w);
  else
    hipLaunchKernelGGL(k_shadow<false>, dim3(g), dim3(PT_BLOCK), 0, st, s, w);
}
void launch_film(hipStream_t st, const RenderParams &rp, const Wave &w, float *film, float *samples,
                 unsigned long long *stats) {
  uint32_t npix = rp.chunk_paths / rp.spp_count;
  hipLaunchKernelGGL(k_film, dim3(blocks_for(npix)), dim3(256), 0, st, rp, w, film, samples, stats);
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

}  // namespace pt
