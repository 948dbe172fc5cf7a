/*
 * or_bsdf.c -- oracle restatement of the reference's materials and PDFs.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Build-time configuration restated: BIASED_SAMPLING (src/CMakeLists.txt:22)
 * and TransportMode::Importance (config.hpp:23).
 *
 * Pinning (tests/test_oracle_bsdf.py): Fresnel KATs of
 * tests/test_dielectric.cpp:80-171 and test_refract.cpp:8-86; Lambert and
 * GGX closed forms (Walter et al. 2007 D and Lambda, Heitz 2018 VNDF);
 * MicrofacetPDF E[1/p] = 2*pi (test_pdf_normalization.cpp:48-76); pdf
 * integrals and sample/pdf consistency by Monte Carlo; alpha -> 0 reduces to
 * the specular cases.  Disney and hair: or_bsdf_disney.c.
 */
#include <complex.h>
#include <string.h>
#include "oracle.h"

/* ---------------- warps (utils/random.hpp, utils/utility.hpp) ---------------- */

/* cosineSampleHemiSphere (random.hpp:25-31) */
static v3 cosine_sample_hemisphere(float r1, float r2) {
  float x = (float)(cos(2.f * M_PI * r1) * sqrtf(r2));
  float y = (float)(sin(2.f * M_PI * r1) * sqrtf(r2));
  float z = sqrtf(1.f - r2);
  return V3(x, y, z);
}

/* ---------------- TrowbridgeReitzDistribution (materialMath.hpp:109-214) ---- */
typedef struct { float ax, ay; } trd_t;

static int trd_smooth(trd_t d) { return fmaxf(d.ax, d.ay) < 1e-3f; }

static float trd_D(trd_t d, v3 wm) {
  float t2 = tan2_theta(wm);
  if (isinf(t2)) return 0.f;
  float c4 = sqr(cos2_theta(wm));
  float e = t2 * (sqr(cos_phi(wm) / d.ax) + sqr(sin_phi(wm) / d.ay));
  return (float)(1.0 / (M_PI * d.ax * d.ay * c4 * sqr(1.f + e)));
}
static float trd_lambda(trd_t d, v3 w) {
  float t2 = tan2_theta(w);
  if (isinf(t2)) return 0.f;
  float a2 = sqr(cos_phi(w) * d.ax) + sqr(sin_phi(w) * d.ay);
  return (sqrtf(1.f + a2 * t2) - 1.f) / 2.f;
}
static float trd_G1(trd_t d, v3 w) { return 1.f / (1.f + trd_lambda(d, w)); }
static float trd_G(trd_t d, v3 wo, v3 wi) { return 1.f / (1.f + trd_lambda(d, wo) + trd_lambda(d, wi)); }
/* D(w, wm) = G1(w)/|cos w| * D(wm) * |w.wm|  (materialMath.hpp:166-174) */
static float trd_Dw(trd_t d, v3 w, v3 wm) {
  return trd_G1(d, w) / abs_cos_theta(w) * trd_D(d, wm) * fabsf(vdot(w, wm));
}
/* sample_wm (materialMath.hpp:178-200): Heitz 2018 visible-normal sampling.
 * vec2::length() of the reference reads an uninitialised SIMD lane
 * (matrix.hpp:457-459); it is taken as zero here. */
static v3 trd_sample_wm(trd_t d, v3 w, float u0, float u1) {
  v3 wh = vnorm(V3(d.ax * w.x, d.ay * w.y, w.z));
  if (wh.z < 0.f) wh = vneg(wh);
  v3 T1 = (wh.z < 0.99999f) ? vnorm(vcross(V3(0, 0, 1), wh)) : V3(1, 0, 0);
  v3 T2 = vcross(wh, T1);
  /* sampleUniformDiskPolar (utility.hpp:593-598) */
  float r = sqrtf(u0);
  float theta = (float)(2 * M_PI * u1);
  float px = (float)(r * cos(theta)), py = (float)(r * sin(theta));
  float h = sqrtf(1.f - sqr(px));
  float t = (1.f + wh.z) / 2.f;
  py = h + t * (py - h); /* gl::lerp(h, p.y, t) */
  float plen = sqrtf(px * px + py * py);
  float pz = sqrtf(fmaxf(0.f, 1.f - sqr(plen)));
  v3 nh = vadd(vadd(vscale(T1, px), vscale(T2, py)), vscale(wh, pz));
  return vnorm(V3(d.ax * nh.x, d.ay * nh.y, fmaxf(1e-6f, nh.z)));
}

/* ---------------- Fresnel for conductors (materialMath.hpp:30-53) ---------- */
static float fresnel_complex1(float ci, float eta_re, float eta_im) {
  float complex eta = eta_re + eta_im * I;
  ci = clampf(ci, 0.f, 1.f);
  float s2i = 1.f - sqr(ci);
  float complex s2t = s2i / (eta * eta);
  float complex ct = csqrtf(1.0f - s2t);
  float complex rpa = (eta * ci - ct) / (eta * ci + ct);
  float complex rpe = (ci - eta * ct) / (ci + eta * ct);
  float na = crealf(rpa) * crealf(rpa) + cimagf(rpa) * cimagf(rpa);
  float ne = crealf(rpe) * crealf(rpe) + cimagf(rpe) * cimagf(rpe);
  return (na + ne) / 2.f;
}
static v3 fresnel_complex(float ci, v3 eta, v3 k) {
  return V3(fresnel_complex1(ci, eta.x, k.x), fresnel_complex1(ci, eta.y, k.y),
            fresnel_complex1(ci, eta.z, k.z));
}

static v3 p3(const pt_material_desc *m, int o) { return V3(m->p[o], m->p[o + 1], m->p[o + 2]); }

/* ---------------- Lambertian (material.hpp:79-137, CosinePDF pdf.hpp:19-37) -- */
static int lambert_scatter(const pt_material_desc *m, const or_hit *h, float u0, float u1,
                           or_srec *s) {
  onb_t b = onb_make(h->n);
  s->flags = OR_DIFFUSE | OR_REFL;
  s->att = vscale(h->has_tex ? h->tex : p3(m, 0), (float)(1.0f / M_PI));
  v3 d = onb_to_world(&b, cosine_sample_hemisphere(u0, u1));
  s->wi = vnorm(d);
  s->pdf = (float)(fmaxf(vdot(vnorm(s->wi), b.w), 0.f) / M_PI);
  return 1;
}
static float cosine_pdf(v3 n, v3 wi) {
  onb_t b = onb_make(n);
  return (float)(fmaxf(vdot(vnorm(wi), b.w), 0.f) / M_PI);
}
/* CosinePDF(n) (pdf.hpp:19-37): get(uc, u) and at(dir), for the integrators
 * that fall back to it when a material has no pdf of its own. */
v3 or_cosine_dir(v3 n, float u0, float u1) {
  onb_t b = onb_make(n);
  return vnorm(onb_to_world(&b, cosine_sample_hemisphere(u0, u1)));
}
float or_cosine_pdf(v3 n, v3 wi) { return cosine_pdf(n, wi); }

/* ---------------- Conductor (conductor.hpp:4-145, mfPDF.hpp:6-62) ---------- */
static float mf_pdf_at(trd_t d, const onb_t *b, v3 wo_l, v3 wi_world) {
  v3 wi = onb_to_local(b, vnorm(wi_world));
  if (wi.z <= 0.f) return 0.f;
  v3 m = vnorm(vadd(wo_l, wi));
  float denom = 4.f * fabsf(vdot(wo_l, m));
  if (denom < 1e-6f) return 0.f;
  return trd_Dw(d, wo_l, m) / denom;
}
static v3 conductor_f(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  trd_t d = {m->p[6], m->p[7]};
  onb_t b = onb_make(h->n);
  v3 wo_l = onb_to_local(&b, vnorm(wo)), wi_l = onb_to_local(&b, vnorm(wi));
  if (wi_l.z <= 0.f || wo_l.z <= 0.f) return vs(0.f);
  if (trd_smooth(d)) return vs(0.f);
  v3 ml = vnorm(vadd(wo_l, wi_l));
  float D = trd_D(d, ml), G = trd_G(d, wo_l, wi_l);
  v3 F = fresnel_complex(fabsf(vdot(wo_l, ml)), p3(m, 0), p3(m, 3));
  float den = 4.f * abs_cos_theta(wo_l) * abs_cos_theta(wi_l);
  return vdivs(vscale(vscale(F, D), G), den);
}
static int conductor_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float u0,
                             float u1, or_srec *s) {
  trd_t d = {m->p[6], m->p[7]};
  v3 wo = vneg(vnorm(ray_d));
  if (trd_smooth(d)) {
    v3 wi = vnorm(pbrt_reflect(wo, h->n));
    if (vdot(h->n, wi) <= 0.f) return 0;
    s->wi = wi;
    s->att = fresnel_complex(fabsf(vdot(h->n, wi)), p3(m, 0), p3(m, 3));
    s->flags = OR_SPECULAR | OR_REFL;
    s->pdf = 0.f;
    return 1;
  }
  onb_t b = onb_make(h->n);
  v3 wo_l = onb_to_local(&b, vnorm(wo));
  v3 ml = trd_sample_wm(d, wo_l, u0, u1);
  v3 wi_l = pbrt_reflect(wo_l, ml);
  if (!same_hemisphere(wo_l, wi_l)) wi_l.z = -wo_l.z; /* BIASED_SAMPLING, mfPDF.hpp:28-30 */
  v3 wi = vnorm(onb_to_world(&b, wi_l));
  if (near_zero(wi)) return 0;
  if (vdot(h->n, wi) <= 0.f) return 0;
  s->wi = wi;
  s->flags = OR_GLOSSY | OR_REFL;
  s->att = conductor_f(m, wo, wi, h);
  s->pdf = mf_pdf_at(d, &b, wo_l, wi);
  return 1;
}
static float conductor_pdf(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  trd_t d = {m->p[6], m->p[7]};
  if (trd_smooth(d)) return 0.f;
  onb_t b = onb_make(h->n);
  v3 wo_l = onb_to_local(&b, vnorm(wo));
  if (vdot(h->n, wi) <= 0.f) return 0.f;
  return mf_pdf_at(d, &b, wo_l, wi);
}

/* ---------------- MFDielectric (dielectric.hpp:5-167, mfDielectricPDF.hpp) -- */
static v3 mfdiel_f(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  float eta = m->p[0];
  trd_t d = {m->p[1], m->p[2]};
  if (eta == 1.f || trd_smooth(d)) return vs(0.f);
  onb_t b = onb_make(h->n);
  v3 wo_l = onb_to_local(&b, vnorm(wo)), wi_l = onb_to_local(&b, vnorm(wi));
  float co = cos_theta(wo_l), ci = cos_theta(wi_l);
  int refl = ci * co > 0.f;
  float etap = 1.f;
  if (!refl) etap = co > 0.f ? eta : (1.f / eta);
  v3 wm = vadd(vscale(wi_l, etap), wo_l);
  if (ci == 0.f || co == 0.f || vlen(wm) == 0.f) return vs(0.f);
  wm = face_forward(vnorm(wm), V3(0, 0, 1));
  if (vdot(wm, wi_l) * ci < 0.f || vdot(wm, wo_l) * co < 0.f) return vs(0.f);
  float F = fresnel_dielectric(vdot(wo_l, wm), eta);
  if (refl) return vs(trd_D(d, wm) * trd_G(d, wo_l, wi_l) * F / fabsf(4.f * ci * co));
  float denom = sqr(vdot(wi_l, wm) + vdot(wo_l, wm) / etap) * ci * co;
  float ft = trd_D(d, wm) * (1.f - F) * trd_G(d, wo_l, wi_l) *
             fabsf(vdot(wi_l, wm) * vdot(wo_l, wm) / denom);
  return vs(ft); /* TransportMode::Importance: no 1/etap^2 (dielectric.hpp:161-162) */
}
/* MFDielectricPDF::at (mfDielectricPDF.hpp:100-146) */
static float mfdiel_pdf_local(float eta, trd_t d, v3 wo, v3 wi) {
  if (eta == 1.f || trd_smooth(d)) return 0.f;
  if (wi.z == 0.f) return 0.f;
  float co = cos_theta(wo), ci = cos_theta(wi);
  int refl = ci * co > 0.f;
  float etap = 1.f;
  if (!refl) etap = co > 0.f ? eta : (1.f / eta);
  v3 wm = vadd(vscale(wi, etap), wo);
  if (ci == 0.f || co == 0.f || vlen(wm) == 0.f) return 0.f;
  wm = face_forward(vnorm(wm), V3(0, 0, 1));
  if (vdot(wm, wi) * ci < 0.f || vdot(wm, wo) * wo.z < 0.f) return 0.f;
  float R = fresnel_dielectric(vdot(wo, wm), eta), T = 1.f - R;
  if (R + T == 0.f) return 0.f;
  if (refl) return trd_Dw(d, wo, wm) / (4.f * fabsf(vdot(wo, wm))) * R / (R + T);
  float denom = sqr(vdot(wi, wm) + vdot(wo, wm) / etap);
  float dwm_dwi = fabsf(vdot(wi, wm)) / denom;
  return trd_Dw(d, wo, wm) * dwm_dwi * T / (R + T);
}
static int mfdiel_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float uc,
                          float u0, float u1, or_srec *s) {
  float eta = m->p[0];
  trd_t d = {m->p[1], m->p[2]};
  v3 wo = vneg(vnorm(ray_d));
  if (eta == 1.f || trd_smooth(d)) {
    float co = vdot(h->n, wo);
    float R = fresnel_dielectric(co, eta), T = 1.f - R;
    if (R == 0.f && T == 0.f) return 0;
    if (uc < R / (R + T)) {
      s->wi = vnorm(pbrt_reflect(wo, h->n));
      s->flags = OR_SPECULAR | OR_REFL;
      s->att = vs(1.f);
      s->pdf = R;
      return 1;
    }
    float etap;
    v3 wi;
    if (!pbrt_refract(wo, h->n, eta, &etap, &wi)) return 0;
    s->wi = wi;
    s->flags = OR_SPECULAR | OR_TRANS;
    s->pdf = T;
    s->att = vs(1.f); /* Importance mode (dielectric.hpp:66-69) */
    return 1;
  }
  onb_t b = onb_make(h->n);
  v3 wo_l = onb_to_local(&b, vnorm(wo));
  /* MFDielectricPDF::get (mfDielectricPDF.hpp:24-98) with BIASED_SAMPLING */
  v3 ml = trd_sample_wm(d, wo_l, u0, u1);
  float R = fresnel_dielectric(vdot(wo_l, ml), eta), T = 1.f - R;
  v3 wi_l;
  if (R + T == 0.f) return 0;
  if (uc < R / (R + T)) {
    wi_l = pbrt_reflect(wo_l, ml);
    if (!same_hemisphere(wo_l, wi_l)) wi_l.z = -wi_l.z;
  } else {
    float es;
    int valid = pbrt_refract(wo_l, ml, eta, &es, &wi_l);
    if (!valid || same_hemisphere(wo_l, wi_l)) wi_l = pbrt_reflect(wo_l, ml);
  }
  v3 wi = onb_to_world(&b, wi_l);
  if (near_zero(wi)) return 0;
  int is_refl = vdot(onb_to_local(&b, wi), onb_to_local(&b, wo)) > 0.f;
  s->wi = wi;
  s->flags = OR_GLOSSY | (is_refl ? OR_REFL : OR_TRANS);
  s->att = mfdiel_f(m, wo, wi, h);
  s->pdf = mfdiel_pdf_local(eta, d, wo_l, onb_to_local(&b, vnorm(wi)));
  return 1;
}
static float mfdiel_pdf(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  float eta = m->p[0];
  trd_t d = {m->p[1], m->p[2]};
  if (eta == 1.f || trd_smooth(d)) return 0.f;
  onb_t b = onb_make(h->n);
  return mfdiel_pdf_local(eta, d, onb_to_local(&b, vnorm(wo)), onb_to_local(&b, vnorm(wi)));
}

/* ---------------- ThinDielectric (thinDielectric.hpp:3-70) ---------------- */
static int thin_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float uc,
                        or_srec *s) {
  float eta = m->p[0];
  v3 wo = vneg(vnorm(ray_d));
  float aco = fabsf(vdot(h->n, wo));
  float R = fresnel_dielectric(aco, eta), T = 1.f - R;
  if (R < 1.f) {
    R += sqr(T) * R / (1.f - sqr(R));
    T = 1.f - R;
  }
  if (R == 0.f && T == 0.f) return 0;
  /* use_split_ray traces both branches weighted R and T (maxdepth_nee.hpp:51-71);
   * choosing one with probability R has the same expectation and is what is
   * done here, as in the GPU path. */
  if (uc < R / (R + T)) {
    s->wi = vnorm(pbrt_reflect(wo, h->n));
    s->flags = OR_SPECULAR | OR_REFL;
    s->att = vs(1.f);
    s->pdf = R;
  } else {
    s->wi = vneg(vnorm(wo));
    s->flags = OR_SPECULAR | OR_TRANS;
    s->att = vs(1.f);
    s->pdf = T;
  }
  return 1;
}

/* ---------------- Phong (material.hpp:303-465, PhongLobePDF pdf.hpp:57-91) -- */
/* gl::reflect (utility.hpp:39-45) */
static v3 gl_reflect(v3 v, v3 n) {
  v3 vv = vnorm(v), nn = vnorm(n);
  float dt = vdot(vv, nn);
  return vnorm(vsub(vv, vscale(nn, 2.f * dt)));
}
static float phong_lobe_pdf(v3 R, float shin, v3 dir) {
  float c = fmaxf(vdot(vnorm(dir), vnorm(R)), 0.f);
  return (float)((shin + 1.0f) / (2.0f * M_PI) * powf(c, shin));
}
static int phong_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float u0,
                         float u1, float aux, or_srec *s) {
  v3 dif = p3(m, 0), spe = p3(m, 3);
  float shin = m->p[9];
  float si = (spe.x + spe.y + spe.z) / 3.f, di = (dif.x + dif.y + dif.z) / 3.f;
  float sp = si / (si + di);
  if (aux < sp) { /* lobe choice from rand_num (material.hpp:336) */
    v3 R = gl_reflect(vnorm(ray_d), h->n);
    v3 Rn = vnorm(R);
    onb_t b = onb_make(R);
    float theta = acosf(powf(u0, 1.f / (shin + 1.f)));
    float phi = (float)(2.0f * M_PI * u1);
    float st = sinf(theta);
    v3 loc = V3(st * cosf(phi), st * sinf(phi), cosf(theta));
    v3 dir = vnorm(onb_to_world(&b, loc));
    s->flags = OR_GLOSSY | OR_REFL;
    s->wi = dir;
    s->att = spe;
    s->pdf = phong_lobe_pdf(Rn, shin, dir);
    return 1;
  }
  onb_t b = onb_make(h->n);
  s->flags = OR_DIFFUSE | OR_REFL;
  s->att = dif;
  s->wi = onb_to_world(&b, cosine_sample_hemisphere(u0, u1));
  s->pdf = 0.f; /* the diffuse branch never sets pdf_val (material.hpp:357-365) */
  return 1;
}
static v3 phong_f(const pt_material_desc *m, v3 wo_w, v3 wi_w, const or_hit *h) {
  v3 dif = p3(m, 0), spe = p3(m, 3);
  float shin = m->p[9];
  v3 wo = vnorm(wo_w), wi = vnorm(wi_w);
  float cti = vdot(h->n, wi);
  if (cti <= 0.f) return vs(0.f);
  float si = (spe.x + spe.y + spe.z) / 3.f, di = (dif.x + dif.y + dif.z) / 3.f;
  float sp = si / (si + di);
  v3 db = vscale(dif, (float)(1.0f / M_PI));
  v3 R = pbrt_reflect(wo, h->n);
  float ca = fmaxf(vdot(wi, R), 0.f);
  float pt = powf(ca, shin);
  float sn = (float)((shin + 2.0f) / (2.0f * M_PI));
  v3 sb = vscale(spe, sn * pt);
  return vadd(vscale(sb, sp), vscale(db, 1.f - sp));
}
static float phong_pdf(const pt_material_desc *m, v3 wo_w, v3 wi_w, const or_hit *h) {
  v3 dif = p3(m, 0), spe = p3(m, 3);
  float shin = m->p[9];
  v3 wo = vnorm(wo_w), wi = vnorm(wi_w);
  float si = (spe.x + spe.y + spe.z) / 3.f, di = (dif.x + dif.y + dif.z) / 3.f;
  float tot = si + di;
  if (tot == 0.f) return 0.f;
  float ps = si / tot, pd = di / tot;
  float pl = 0.f, pdl = 0.f;
  if (ps > 0.f) pl = phong_lobe_pdf(gl_reflect(vneg(wo), vnorm(h->n)), shin, wi);
  if (pd > 0.f) {
    float c = vdot(wi, vnorm(h->n));
    if (c > 0.f) pdl = (float)(c / M_PI);
  }
  return ps * pl + pd * pdl;
}

/* ---------------- Kajiya-Kay (no reference code: README.md:74 lists it,
 * TODO.md:20 is unchecked).  The Kajiya & Kay (1989) fibre model about the
 * fibre tangent t: the hit's hair tangent on curves (HitRecord::hair_tangent,
 * curve.hpp:101-129); on a surface the first axis u of OrthoBasis(n), a
 * direction in the tangent plane (a hair card):
 *   f = kd/pi sin(t,wi)
 *     + ks (p+2)/(2pi) max(0, sin(t,wi) sin(t,wo) - cos(t,wi) cos(t,wo))^p
 * The diffuse term is Lambert integrated over the lit half of a thin
 * cylinder.  The specular term is cos^p of the angle between wo and the cone
 * of mirror directions: light reflected by a cylinder keeps its component
 * along the axis, so the cone is t.wo = -t.wi.  Sampled by the cosine lobe
 * about n.  Pinned by tests/test_oracle_bsdf.py: the closed-form values on
 * the cone, at normal incidence and retro-reflection, the tangent choice, the
 * pdf's normalisation and sample/pdf consistency. */
static v3 kk_tangent(const or_hit *h) {
  if (!(fabsf(h->tangent.x) < 1e-6f && fabsf(h->tangent.y) < 1e-6f && fabsf(h->tangent.z) < 1e-6f))
    return vnorm(h->tangent);
  onb_t b = onb_make(h->n);
  return b.u;
}
static v3 kk_f(const pt_material_desc *m, v3 wo_w, v3 wi_w, const or_hit *h) {
  v3 t = kk_tangent(h), wo = vnorm(wo_w), wi = vnorm(wi_w);
  float ci = vdot(t, wi), co = vdot(t, wo);
  float si = sqrtf(fmaxf(0.f, 1.f - ci * ci)), so = sqrtf(fmaxf(0.f, 1.f - co * co));
  float ex = m->p[6];
  float spec = powf(fmaxf(0.f, si * so - ci * co), ex) * (float)((ex + 2.f) / (2.f * M_PI));
  v3 kd = vscale(p3(m, 0), (float)(si / M_PI));
  return vadd(kd, vscale(p3(m, 3), spec));
}
static float kk_pdf(v3 wo, v3 wi, const or_hit *h) {
  (void)wo;
  return cosine_pdf(h->n, wi);
}
static int kk_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float u0,
                      float u1, or_srec *s) {
  onb_t b = onb_make(h->n);
  v3 wo = vneg(vnorm(ray_d));
  s->wi = vnorm(onb_to_world(&b, cosine_sample_hemisphere(u0, u1)));
  s->pdf = cosine_pdf(h->n, s->wi);
  if (s->pdf <= 0.f) return 0;
  s->flags = OR_GLOSSY | OR_REFL;
  s->att = kk_f(m, wo, s->wi, h);
  return 1;
}

/* Disney and hair: same translation unit, so they share the helpers above */
int or_disney_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float uc,
                      float u0, float u1, or_srec *s);
v3 or_disney_f(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h);
float or_disney_pdf(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h);
int or_hair_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float uc,
                    float u0, float u1, or_srec *s);
v3 or_hair_f(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h);
float or_hair_pdf(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h);
#include "or_bsdf_disney.c"

/* ---------------- dispatch (Material virtuals, material.hpp:15-77) -------- */
int or_is_emitter(const pt_material_desc *m) { return m->kind == PT_MAT_EMITTER; }

int or_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float uc, float u0,
               float u1, float aux, or_srec *s) {
  memset(s, 0, sizeof(*s));
  switch (m->kind) {
  case PT_MAT_LAMBERTIAN: return lambert_scatter(m, h, u0, u1, s);
  case PT_MAT_CONDUCTOR: return conductor_scatter(m, ray_d, h, u0, u1, s);
  case PT_MAT_MFDIELECTRIC: return mfdiel_scatter(m, ray_d, h, uc, u0, u1, s);
  case PT_MAT_THINDIELECTRIC: return thin_scatter(m, ray_d, h, uc, s);
  case PT_MAT_PHONG: return phong_scatter(m, ray_d, h, u0, u1, aux, s);
  case PT_MAT_DISNEY: return or_disney_scatter(m, ray_d, h, uc, u0, u1, s);
  case PT_MAT_HAIR: return or_hair_scatter(m, ray_d, h, uc, u0, u1, s);
  case PT_MAT_KAJIYA_KAY: return kk_scatter(m, ray_d, h, u0, u1, s);
  default: return 0; /* emitters and unknown kinds do not scatter */
  }
}

v3 or_bsdf_f(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  switch (m->kind) {
  case PT_MAT_LAMBERTIAN: return vscale(h->has_tex ? h->tex : p3(m, 0), (float)(1.0f / M_PI));
  case PT_MAT_CONDUCTOR: return conductor_f(m, wo, wi, h);
  case PT_MAT_MFDIELECTRIC: return mfdiel_f(m, wo, wi, h);
  case PT_MAT_PHONG: return phong_f(m, wo, wi, h);
  case PT_MAT_DISNEY: return or_disney_f(m, wo, wi, h);
  case PT_MAT_HAIR: return or_hair_f(m, wo, wi, h);
  case PT_MAT_KAJIYA_KAY: return kk_f(m, wo, wi, h);
  default: return vs(0.f);
  }
}

/* Material::scatter_pdf(wo, wi, rec) */
float or_bsdf_pdf(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  switch (m->kind) {
  case PT_MAT_LAMBERTIAN: return cosine_pdf(h->n, wi);
  case PT_MAT_CONDUCTOR: return conductor_pdf(m, wo, wi, h);
  case PT_MAT_MFDIELECTRIC: return mfdiel_pdf(m, wo, wi, h);
  case PT_MAT_PHONG: return phong_pdf(m, wo, wi, h);
  case PT_MAT_DISNEY: return or_disney_pdf(m, wo, wi, h);
  case PT_MAT_HAIR: return or_hair_pdf(m, wo, wi, h);
  case PT_MAT_KAJIYA_KAY: return kk_pdf(wo, wi, h);
  default: return 0.f;
  }
}

/* DiffuseEmitter::emit (material.hpp:487-493) */
v3 or_emit(const pt_material_desc *m, v3 ray_d, const or_hit *h) {
  if (m->kind != PT_MAT_EMITTER) return vs(0.f);
  if (vdot(h->n, vneg(ray_d)) > 0.f) return vscale(p3(m, 0), m->p[3]);
  return vs(0.f);
}

int or_bsdf_eval_batch(const pt_material_desc *m, const float *wo, const float *wi,
                       const float *nrm, const uint8_t *inside, size_t n, float *f,
                       float *pdf) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    or_hit h;
    memset(&h, 0, sizeof(h));
    h.n = V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
    h.inside = inside ? inside[i] : 1;
    v3 o = V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    v3 w = V3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
    v3 r = or_bsdf_f(m, o, w, &h);
    f[3 * i] = r.x; f[3 * i + 1] = r.y; f[3 * i + 2] = r.z;
    pdf[i] = or_bsdf_pdf(m, o, w, &h);
  }
  return 0;
}

int or_bsdf_sample_batch(const pt_material_desc *m, const float *wo, const float *nrm,
                         const uint8_t *inside, const float *uc, const float *u2,
                         const float *aux, size_t n, float *wi, float *f, float *pdf,
                         uint32_t *flags, uint8_t *ok) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    or_hit h;
    memset(&h, 0, sizeof(h));
    h.n = V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
    h.inside = inside ? inside[i] : 1;
    v3 o = V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    or_srec s;
    int r = or_scatter(m, vneg(o), &h, uc[i], u2[2 * i], u2[2 * i + 1], aux ? aux[i] : 0.5f, &s);
    ok[i] = (uint8_t)r;
    wi[3 * i] = s.wi.x; wi[3 * i + 1] = s.wi.y; wi[3 * i + 2] = s.wi.z;
    f[3 * i] = s.att.x; f[3 * i + 1] = s.att.y; f[3 * i + 2] = s.att.z;
    pdf[i] = s.pdf;
    flags[i] = s.flags;
  }
  return 0;
}

/* ---------------- exported helpers for closed-form tests ---------------- */
float or_fresnel_dielectric(float ci, float eta) { return fresnel_dielectric(ci, eta); }
void or_fresnel_complex(float ci, const float *eta, const float *k, float *out) {
  v3 r = fresnel_complex(ci, V3(eta[0], eta[1], eta[2]), V3(k[0], k[1], k[2]));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
float or_trd_D(float ax, float ay, const float *wm) {
  trd_t d = {ax, ay};
  return trd_D(d, V3(wm[0], wm[1], wm[2]));
}
float or_trd_Dw(float ax, float ay, const float *w, const float *wm) {
  trd_t d = {ax, ay};
  return trd_Dw(d, V3(w[0], w[1], w[2]), V3(wm[0], wm[1], wm[2]));
}
float or_trd_G1(float ax, float ay, const float *w) {
  trd_t d = {ax, ay};
  return trd_G1(d, V3(w[0], w[1], w[2]));
}
void or_trd_sample_wm(float ax, float ay, const float *w, float u0, float u1, float *out) {
  trd_t d = {ax, ay};
  v3 m = trd_sample_wm(d, V3(w[0], w[1], w[2]), u0, u1);
  out[0] = m.x; out[1] = m.y; out[2] = m.z;
}
int or_refract(const float *wi, const float *n, float eta, float *wt, float *etap) {
  v3 t = vs(0.f);
  int ok = pbrt_refract(V3(wi[0], wi[1], wi[2]), V3(n[0], n[1], n[2]), eta, etap, &t);
  wt[0] = t.x; wt[1] = t.y; wt[2] = t.z;
  return ok;
}
