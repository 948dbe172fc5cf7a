/*
 * or_bsdf_disney.c -- oracle restatement of DisneyPrincipledBSDF and
 * HairMarschner.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Compiled as part of or_bsdf.c (included there) so that it shares the
 * microfacet and Fresnel helpers.
 *
 * Reference behaviour kept as is:
 *   - HitRecord::is_inside is the front-face flag; on front faces the Disney
 *     material evaluates and samples its glass lobe only
 *     (disneyPrincipledBSDF.cpp:13-18, disneyPrincipledBSDF.hpp:78-79).
 *   - the metal lobe of the mixture is f_variant with its default arguments,
 *     C0 = 0.5 * baseColor (disneyPrincipledBSDF.hpp:85, disneyMetal.hpp:100).
 *   - the chosen lobe samples with u = (uc_remapped, u.y); specular lobes
 *     return without the lobe pmf; sheen is left out of the pdf
 *     (disneyPrincipledBSDF.cpp:36-73, 101-113).
 *   - hair: evalMarschner feeds the unrotated theta_o to Mp
 *     (hairMarsch.cpp:67-97); HairPDF::at sums p < pMax only (hairPDF.cpp:111);
 *     fastExp is the reference's polynomial (utility.hpp:248-268).
 *
 * Pinning (tests/test_oracle_disney_hair.py): lobe closed forms (Burley 2012
 * diffuse retro-reflection at grazing, Schlick), the GTR1 clearcoat
 * distribution integrating to 1, mixture-weight identities, the hair pdf
 * integrating to 1 (tests/test_hair_pdf.cpp:6-39, test_pdf_normalization.cpp:
 * 78-116, tolerance 0.02 as there) and sample/pdf consistency.  The exact
 * values of the quirky mixture (A-7, A-8) have no closed form: parity
 * unpinned beyond these invariants (no reference run is possible here).
 */

/* colors.hpp:42-46 */
static float rgb_to_luminance(v3 c) { return vdot(c, V3(0.2126f, 0.7152f, 0.0722f)); }

/* fresnelSchlick (materialMath.hpp:55-58): pow in double */
static v3 fresnel_schlick(float c, v3 F0) {
  float w = (float)pow(1.0 - c, 5.0);
  return V3(F0.x + (1.f - F0.x) * w, F0.y + (1.f - F0.y) * w, F0.z + (1.f - F0.z) * w);
}
/* hackedSchlick (materialMath.hpp:60-63) */
static float hacked_schlick(float c, float F90) { return 1.f + (F90 - 1.f) * powf(1.f - c, 5.f); }

/* TrowbridgeReitzDistribution::fromRoughnessAnisotropic (materialMath.hpp:119-132) */
static trd_t trd_from_roughness(float roughness, float aniso) {
  aniso = fmaxf(0.0f, fminf(aniso, 1.0f / 0.9f - 0.00001f));
  float aspect = safe_sqrt(1.0f - 0.9f * aniso);
  if (aspect == 0.0f) aspect = 0.0001f;
  trd_t d;
  d.ax = fmaxf(0.0001f, sqr(roughness) / aspect);
  d.ay = fmaxf(0.0001f, sqr(roughness) * aspect);
  return d;
}

typedef struct {
  v3 base;
  float spec_trans, metallic, subsurface, specular, roughness, specular_tint, anisotropic, sheen,
      sheen_tint, clearcoat, clearcoat_gloss, eta;
  float dw, mw, gw, cw, sw;
  trd_t dist;
  float cc_alpha;
} disney_t;

static disney_t disney_params(const pt_material_desc *m, const or_hit *h) {
  disney_t d;
  d.base = h->has_tex ? h->tex : V3(m->p[0], m->p[1], m->p[2]);
  d.spec_trans = m->p[3]; d.metallic = m->p[4]; d.subsurface = m->p[5]; d.specular = m->p[6];
  d.roughness = m->p[7]; d.specular_tint = m->p[8]; d.anisotropic = m->p[9]; d.sheen = m->p[10];
  d.sheen_tint = m->p[11]; d.clearcoat = m->p[12]; d.clearcoat_gloss = m->p[13]; d.eta = m->p[14];
  /* disneyPrincipledBSDF.hpp:57-62 */
  d.dw = (1 - d.metallic) * (1 - d.spec_trans);
  d.mw = (1 - d.spec_trans * (1 - d.metallic));
  d.gw = (1 - d.metallic) * d.spec_trans;
  d.cw = (float)(0.25 * d.clearcoat);
  d.sw = (1 - d.metallic) * d.sheen;
  d.dist = trd_from_roughness(d.roughness, d.anisotropic);
  /* ClearcoatTRD::fromClearcoatGloss (materialMath.hpp:224-228) */
  d.cc_alpha = (1.f - d.clearcoat_gloss) * 0.1f + d.clearcoat_gloss * 0.001f;
  return d;
}

/* ---- DisneyDiffuse (disneyDiffuse.hpp:8-88) ---- */
static v3 ddiffuse_f(const disney_t *d, v3 wo_w, v3 wi_w, const or_hit *h) {
  onb_t b = onb_make(h->n);
  v3 wo = onb_to_local(&b, wo_w), wi = onb_to_local(&b, wi_w);
  float aci = fabsf(wi.z), aco = fabsf(wo.z);
  v3 wh = vnorm(vadd(wo, wi));
  float whwi = vdot(wh, wi);
  float fd90 = 0.5f + 2.f * d->roughness * powf(fabsf(whwi), 2.f);
  v3 base_diff = vscale(d->base, (float)(1.f / M_PI));
  base_diff = vscale(base_diff, hacked_schlick(aci, fd90) * hacked_schlick(aco, fd90));
  float fss90 = d->roughness * powf(fabsf(whwi), 2.f);
  v3 base_ss = vscale(vscale(d->base, 1.25f), (float)(1.f / M_PI));
  base_ss = vscale(base_ss, hacked_schlick(aci, fss90) * hacked_schlick(aco, fss90) *
                                (1.f / (aci + aco) - 0.5f) + 0.5f);
  return vadd(vscale(base_diff, 1.f - d->subsurface), vscale(base_ss, d->subsurface));
}

/* ---- DisneyMetal (disneyMetal.hpp:9-140) ---- */
static v3 dmetal_f_c0(const disney_t *d, v3 wo_w, v3 wi_w, const or_hit *h, v3 c0) {
  onb_t b = onb_make(h->n);
  v3 wo = onb_to_local(&b, wo_w), wi = onb_to_local(&b, wi_w);
  v3 wh = vnorm(vadd(wo, wi));
  if (wi.z <= 0.f || wo.z <= 0.f || trd_smooth(d->dist)) return vs(0.f);
  float dwhwi = fabsf(vdot(wh, wi));
  v3 F = fresnel_schlick(dwhwi, c0);
  float D = trd_D(d->dist, wh), G = trd_G(d->dist, wo, wi);
  return vdivs(vscale(vscale(F, D), G), fabsf(4.f * wi.z * wo.z));
}
static v3 dmetal_f(const disney_t *d, v3 wo, v3 wi, const or_hit *h) { return dmetal_f_c0(d, wo, wi, h, d->base); }
/* f_variant with default specular_tint .5, eta 1, specular .5, metallic .5:
 * Ks = .5 + .5 base, R0 = etaToR0(1) = 0, C0 = .5*0*(1-.5)*Ks + .5*base. */
static v3 dmetal_f_variant(const disney_t *d, v3 wo, v3 wi, const or_hit *h) {
  v3 ks = vadd(vs(1.f - 0.5f), vscale(d->base, 0.5f));
  float r0 = sqr((1.f - 1.f) / (1.f + 1.f));
  v3 c0 = vadd(vscale(vscale(ks, r0 * (1.f - 0.5f)), 0.5f), vscale(d->base, 0.5f));
  return dmetal_f_c0(d, wo, wi, h, c0);
}
static float dmetal_pdf(const disney_t *d, v3 wo, v3 wi, const or_hit *h) {
  if (trd_smooth(d->dist)) return 0.f;
  onb_t b = onb_make(h->n);
  return mf_pdf_at(d->dist, &b, onb_to_local(&b, vnorm(wo)), wi);
}
static int dmetal_scatter(const disney_t *d, v3 wo, const or_hit *h, float u0, float u1, or_srec *s) {
  if (trd_smooth(d->dist)) {
    v3 wi = vnorm(pbrt_reflect(wo, h->n));
    if (vdot(h->n, wi) <= 0.f) return 0;
    s->wi = wi;
    s->att = fresnel_schlick(fabsf(vdot(h->n, wi)), d->base);
    s->flags = OR_SPECULAR | OR_REFL;
    s->pdf = 0.f;
    return 1;
  }
  onb_t b = onb_make(h->n);
  v3 wo_l = onb_to_local(&b, vnorm(wo));
  v3 ml = trd_sample_wm(d->dist, wo_l, u0, u1);
  v3 wi_l = pbrt_reflect(wo_l, ml);
  if (!same_hemisphere(wo_l, wi_l)) wi_l.z = -wo_l.z;
  v3 wi = vnorm(onb_to_world(&b, wi_l));
  if (near_zero(wi)) return 0;
  if (vdot(h->n, wi) <= 0.f) return 0;
  s->wi = vnorm(wi);
  s->flags = OR_GLOSSY | OR_REFL;
  s->att = dmetal_f(d, wo, wi, h);
  s->pdf = mf_pdf_at(d->dist, &b, wo_l, wi);
  return 1;
}

/* ---- DisneyGlass (disneyGlass.hpp:5-168) ---- */
static v3 dglass_f(const disney_t *d, v3 wo, v3 wi, const or_hit *h) {
  float eta = d->eta;
  trd_t t = d->dist;
  if (eta == 1.f || trd_smooth(t)) return vs(0.f);
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
  if (refl) return vscale(d->base, trd_D(t, wm) * trd_G(t, wo_l, wi_l) * F / fabsf(4.f * ci * co));
  float den = sqr(vdot(wi_l, wm) + vdot(wo_l, wm) / etap) * ci * co;
  v3 sc = V3(safe_sqrt(d->base.x), safe_sqrt(d->base.y), safe_sqrt(d->base.z));
  float k = trd_D(t, wm) * (1.f - F) * trd_G(t, wo_l, wi_l) * fabsf(vdot(wi_l, wm) * vdot(wo_l, wm) / den);
  return vscale(sc, k);
}
static float dglass_pdf(const disney_t *d, v3 wo, v3 wi, const or_hit *h) {
  if (d->eta == 1.f || trd_smooth(d->dist)) return 0.f;
  onb_t b = onb_make(h->n);
  return mfdiel_pdf_local(d->eta, d->dist, onb_to_local(&b, vnorm(wo)), onb_to_local(&b, vnorm(wi)));
}
static int dglass_scatter(const disney_t *d, v3 wo, const or_hit *h, float uc, float u0, float u1,
                          or_srec *s) {
  float eta = d->eta;
  trd_t t = d->dist;
  if (eta == 1.f || trd_smooth(t)) {
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
    s->att = vs(1.f);
    s->pdf = T;
    return 1;
  }
  onb_t b = onb_make(h->n);
  v3 wo_l = onb_to_local(&b, vnorm(wo));
  v3 ml = trd_sample_wm(t, wo_l, u0, u1);
  float R = fresnel_dielectric(vdot(wo_l, ml), eta), T = 1.f - R;
  if (R + T == 0.f) return 0;
  v3 wi_l;
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
  s->att = dglass_f(d, wo, wi, h);
  s->pdf = mfdiel_pdf_local(eta, t, wo_l, onb_to_local(&b, vnorm(wi)));
  return 1;
}

/* ---- DisneyClearcoat (disneyClearcoat.hpp:7-79, ClearcoatPDF clearcoatPDF.hpp:6-61,
 * ClearcoatTRD materialMath.hpp:216-279) ---- */
static float cc_D(float a, v3 hh) {
  float upper = a * a - 1;
  float lower = (float)(M_PI * logf(sqr(a)));
  float term = 1 + (sqr(a) - 1) * sqr(hh.z);
  return upper / (lower * term);
}
static float cc_lambda(v3 w) {
  float t2 = tan2_theta(w);
  if (isinf(t2)) return 0.f;
  double ca = cos_phi(w) * 0.25, sa = sin_phi(w) * 0.25; /* materialMath.hpp:257-258, in double */
  float a2 = (float)(ca * ca + sa * sa);
  return (sqrtf(1 + a2 * t2) - 1) / 2;
}
static float cc_G(v3 wo, v3 wi) { return 1.f / (1.f + cc_lambda(wo) + cc_lambda(wi)); }
static v3 dcc_f(const disney_t *d, v3 wo_w, v3 wi_w, const or_hit *h) {
  onb_t b = onb_make(h->n);
  v3 wo = onb_to_local(&b, wo_w), wi = onb_to_local(&b, wi_w);
  v3 wh = vnorm(vadd(wo, wi));
  if (wi.z <= 0.f || wo.z <= 0.f) return vs(0.f);
  float a = fabsf(vdot(wh, wi));
  float r0 = sqr((1.5f - 1) / (1.5f + 1));
  v3 Fc = fresnel_schlick(a, vs(r0));
  float Dc = cc_D(d->cc_alpha, wh), Gc = cc_G(wo, wi);
  return vdivs(vscale(vscale(Fc, Dc), Gc), fabsf(4.f * wi.z * wo.z));
}
static float cc_pdf_at(float alpha, const onb_t *b, v3 wo_l, v3 wi_w) {
  v3 wi = onb_to_local(b, vnorm(wi_w));
  if (wi.z <= 0.f) return 0.f;
  v3 hh = vnorm(vadd(wo_l, wi));
  float den = 4.f * fabsf(vdot(wi, hh));
  if (den < 1e-6f) return 0.f;
  return cc_D(alpha, hh) * fabsf(hh.z) / den;
}
static float dcc_pdf(const disney_t *d, v3 wo, v3 wi, const or_hit *h) {
  onb_t b = onb_make(h->n);
  return cc_pdf_at(d->cc_alpha, &b, onb_to_local(&b, vnorm(wo)), wi);
}
static int dcc_scatter(const disney_t *d, v3 wo, const or_hit *h, float u0, float u1, or_srec *s) {
  onb_t b = onb_make(h->n);
  v3 wo_l = onb_to_local(&b, vnorm(wo));
  float a = d->cc_alpha;
  /* ClearcoatTRD::sample_h (materialMath.hpp:262-278) */
  float upper = 1 - powf(sqr(a), 1 - u0);
  float lower = 1 - sqr(a);
  float ch = safe_sqrt(upper / lower);
  float sh = safe_sqrt(1 - sqr(ch));
  float az = (float)(2 * M_PI * u1);
  v3 hl = V3((float)(sh * cos(az)), (float)(sh * sin(az)), ch);
  v3 wi_l = pbrt_reflect(wo_l, hl);
  if (!same_hemisphere(wo_l, wi_l)) wi_l.z = -wo_l.z; /* BIASED_SAMPLING (clearcoatPDF.hpp:27-29) */
  v3 wi = vnorm(onb_to_world(&b, wi_l));
  if (vdot(h->n, wi) <= 0.f) return 0;
  s->wi = vnorm(wi);
  s->flags = OR_GLOSSY | OR_REFL;
  s->att = dcc_f(d, wo, wi, h);
  s->pdf = cc_pdf_at(a, &b, wo_l, wi);
  return 1;
}

/* ---- DisneySheen (disneySheen.hpp:7-72) ---- */
static v3 dsheen_f(const disney_t *d, v3 wo_w, v3 wi_w, const or_hit *h) {
  onb_t b = onb_make(h->n);
  v3 wo = onb_to_local(&b, wo_w), wi = onb_to_local(&b, wi_w);
  v3 wh = vnorm(vadd(wo, wi));
  float a = fabsf(vdot(wh, wi));
  float lum = rgb_to_luminance(d->base);
  v3 ctint = lum > 0.0f ? vdivs(d->base, lum) : vs(1.0f);
  v3 cs = vadd(vs(1.f), vscale(vsub(ctint, vs(1.f)), d->sheen_tint)); /* gl::lerp(1, Ctint, tint) */
  return vscale(cs, powf(1 - a, 5.f));
}

/* ---- DisneyPrincipledBSDF (disneyPrincipledBSDF.hpp:72-99, .cpp:3-114) ---- */
v3 or_disney_f(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  disney_t d = disney_params(m, h);
  if (h->inside) return dglass_f(&d, wo, wi, h);
  v3 r = vs(0.f);
  r = vadd(r, vscale(ddiffuse_f(&d, wo, wi, h), d.dw));
  r = vadd(r, vscale(dmetal_f_variant(&d, wo, wi, h), d.mw));
  r = vadd(r, vscale(dglass_f(&d, wo, wi, h), d.gw));
  r = vadd(r, vscale(dcc_f(&d, wo, wi, h), d.cw));
  r = vadd(r, vscale(dsheen_f(&d, wo, wi, h), d.sw));
  return r;
}

float or_disney_pdf(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  disney_t d = disney_params(m, h);
  if (h->inside) return dglass_pdf(&d, wo, wi, h);
  float sum = d.dw + d.mw + d.gw + d.cw;
  float p = 0.f;
  p += (d.dw / sum) * cosine_pdf(h->n, wi);
  p += (d.mw / sum) * dmetal_pdf(&d, wo, wi, h);
  p += (d.gw / sum) * dglass_pdf(&d, wo, wi, h);
  p += (d.cw / sum) * dcc_pdf(&d, wo, wi, h);
  return p;
}

/* gl::sampleDiscrete (utility.hpp:155-190) */
static int sample_discrete(const float *w, int n, float u, float *pmf, float *u_remap) {
  float sum = 0.f;
  for (int i = 0; i < n; ++i) sum += w[i];
  float up = u * sum;
  if (up == sum) up = nextafterf(up, -INFINITY);
  int off = 0;
  float acc = 0.f;
  while (off < n - 1 && acc + w[off] <= up) {
    acc += w[off];
    ++off;
  }
  if (pmf) *pmf = w[off] / sum;
  if (u_remap) {
    float r = (up - acc) / w[off];
    *u_remap = r < 0x1.fffffep-1f ? r : 0x1.fffffep-1f;
  }
  return off;
}

int or_disney_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float uc, float u0,
                      float u1, or_srec *s) {
  disney_t d = disney_params(m, h);
  v3 wo = vneg(vnorm(ray_d));
  if (h->inside) return dglass_scatter(&d, wo, h, uc, u0, u1, s);
  float w[4] = {d.dw, d.mw, d.gw, d.cw};
  float pmf, ucr;
  int lobe = sample_discrete(w, 4, uc, &pmf, &ucr);
  if (lobe < 0 || pmf <= 1e-6f) return 0;
  int ok;
  /* the chosen lobe sees u = (uc_remapped, u.y) (disneyPrincipledBSDF.cpp:47-48) */
  float a0 = ucr, a1 = u1;
  (void)u0;
  switch (lobe) {
  case 0: { /* DisneyDiffuse::scatter (disneyDiffuse.hpp:21-38) */
    onb_t b = onb_make(h->n);
    v3 wiw = onb_to_world(&b, cosine_sample_hemisphere(a0, a1));
    s->wi = vnorm(wiw);
    s->pdf = cosine_pdf(h->n, wiw);
    s->flags = OR_DIFFUSE | OR_REFL;
    s->att = ddiffuse_f(&d, wo, wiw, h);
    ok = 1;
    break;
  }
  case 1: ok = dmetal_scatter(&d, wo, h, a0, a1, s); break;
  case 2: ok = dglass_scatter(&d, wo, h, ucr, a0, a1, s); break;
  default: ok = dcc_scatter(&d, wo, h, a0, a1, s); break;
  }
  if (!ok) return 0;
  if (s->flags & OR_SPECULAR) return 1;
  if (s->pdf <= 1e-6f) return 0;
  v3 wiw = vnorm(s->wi);
  s->att = or_disney_f(m, wo, wiw, h);
  s->pdf = or_disney_pdf(m, wo, wiw, h);
  if (s->pdf <= 1e-7f) return 0;
  return 1;
}

/* ======================= HairMarschner ======================= */
#define HAIR_PMAX 3
typedef struct {
  v3 sigma_a;
  float h, eta, beta_m, beta_n;
  float v[HAIR_PMAX + 1];
  float s;
  float sin2k[HAIR_PMAX], cos2k[HAIR_PMAX];
} hair_t;

/* HairMarschner::HairMarschner (hairMarsch.hpp:5-29) */
static hair_t hair_params(const pt_material_desc *m) {
  hair_t H;
  H.sigma_a = V3(m->p[0], m->p[1], m->p[2]);
  H.h = m->p[3];
  H.eta = m->p[4];
  H.beta_m = m->p[5];
  H.beta_n = m->p[6];
  float alpha = m->p[7];
  float bm = H.beta_m, bn = H.beta_n;
  H.v[0] = sqr(0.726f * bm + 0.812f * sqr(bm) + 3.7f * powf(bm, 20.f));
  H.v[1] = (float)(.25 * H.v[0]);
  H.v[2] = 4 * H.v[0];
  for (int p = 3; p <= HAIR_PMAX; ++p) H.v[p] = H.v[2];
  const float sqrt_pi_over8 = 0.626657069f;
  H.s = sqrt_pi_over8 * (0.265f * bn + 1.194f * sqr(bn) + 5.372f * powf(bn, 22.f));
  H.sin2k[0] = sinf(alpha);
  H.cos2k[0] = safe_sqrt(1 - sqr(H.sin2k[0]));
  for (int i = 1; i < HAIR_PMAX; ++i) {
    H.sin2k[i] = 2 * H.cos2k[i - 1] * H.sin2k[i - 1];
    H.cos2k[i] = sqr(H.cos2k[i - 1]) - sqr(H.sin2k[i - 1]);
  }
  return H;
}

/* gl::I0f, logI0f, fastExp (utility.hpp:213-268).  fastExp's exponent check
 * throws in the reference when the result would exceed FLT_MAX; here the
 * value saturates to +inf. */
static float i0f(float x) {
  float term = 1.0f, sum = term, x2 = x * x;
  for (int k = 1; k < 500; ++k) {
    term *= x2 / (4.0f * k * k);
    sum += term;
    if (term < 1e-6f * sum) break;
    if (sum > FLT_MAX * 0.5f) break;
  }
  return sum;
}
static float logi0f(float x) {
  if (x > 12.0f) return x - 0.5f * logf(2.0f * 3.14159265358979323846f * x);
  return logf(i0f(x));
}
static float fast_exp(float x) {
  float xp = x * 1.442695041f;
  float fxp = floorf(xp);
  float f = xp - fxp;
  int i = (int)fxp;
  float two_to_f = fmaf(f, fmaf(f, fmaf(f, 0.0781455737f, 0.226173572f), 0.695556856f), 1.f);
  uint32_t bits;
  memcpy(&bits, &two_to_f, 4);
  int e = (int)(bits >> 23) - 127 + i;
  if (e < -126) return 0.f;
  if (e > 127) return INFINITY;
  bits &= 0x807FFFFFu;
  bits |= (uint32_t)(e + 127) << 23;
  float r;
  memcpy(&r, &bits, 4);
  return r;
}
/* logistic helpers (utility.hpp:275-310) */
static float logistic(float x, float s) {
  x = fabsf(x);
  return expf(-x / s) / (s * sqr(1 + expf(-x / s)));
}
static float logistic_cdf(float x, float s) { return 1 / (1 + expf(-x / s)); }
static float trimmed_logistic(float x, float s, float a, float b) {
  return logistic(x, s) / (logistic_cdf(b, s) - logistic_cdf(a, s));
}
static float sample_trimmed_logistic(float u, float s, float a, float b) {
  float pa = 1 / (1 + expf(-a / s)), pb = 1 / (1 + expf(-b / s));
  u = pa + u * (pb - pa);
  float x = -s * logf(1 / u - 1);
  return clampf(x, a, b);
}

static float hair_Mp(float ci, float co, float si, float so, float v) {
  float a = ci * co / v, b = si * so / v;
  if ((double)v <= .1) return fast_exp(logi0f(a) - b - 1 / v + 0.6931f + logf(1 / (2 * v)));
  return (fast_exp(-b) * i0f(a)) / (sinhf(1 / v) * 2 * v);
}
static float hair_phi(int p, float go, float gt) { return (float)(2 * p * gt - 2 * go + p * M_PI); }
static float hair_Np(float phi, int p, float s, float go, float gt) {
  float dphi = phi - hair_phi(p, go, gt);
  while (dphi > M_PI) dphi = (float)(dphi - 2 * M_PI);
  while (dphi < -M_PI) dphi = (float)(dphi + 2 * M_PI);
  return trimmed_logistic(dphi, s, (float)-M_PI, (float)M_PI);
}
/* HairMarschner::Ap (hairMarsch.hpp:95-112) */
static void hair_Ap(float co, float eta, float h, v3 T, v3 ap[HAIR_PMAX + 1]) {
  float cgo = safe_sqrt(1 - sqr(h));
  float c = co * cgo;
  float f = fresnel_dielectric(c, eta);
  ap[0] = vs(f);
  ap[1] = vscale(T, sqr(1 - f));
  for (int p = 2; p < HAIR_PMAX; ++p) ap[p] = vscale(vmul(ap[p - 1], T), f);
  v3 den = vsub(vs(1.f), vscale(T, f));
  if (!near_zero(den))
    ap[HAIR_PMAX] = V3(ap[HAIR_PMAX - 1].x * f * T.x / den.x, ap[HAIR_PMAX - 1].y * f * T.y / den.y,
                       ap[HAIR_PMAX - 1].z * f * T.z / den.z);
  else
    ap[HAIR_PMAX] = vs(0.f);
}
/* HairMarschner::ApPDF (hairMarsch.hpp:127-157) */
static void hair_ap_pdf(const hair_t *H, float co, float out[HAIR_PMAX + 1]) {
  float so = safe_sqrt(1 - sqr(co));
  float st = so / H->eta;
  float ct = safe_sqrt(1 - sqr(st));
  float etap = safe_sqrt(sqr(H->eta) - sqr(so)) / co;
  float sgt = H->h / etap;
  float cgt = safe_sqrt(1 - sqr(sgt));
  v3 T = V3(expf(-H->sigma_a.x * (2 * cgt / ct)), expf(-H->sigma_a.y * (2 * cgt / ct)),
            expf(-H->sigma_a.z * (2 * cgt / ct)));
  v3 ap[HAIR_PMAX + 1];
  hair_Ap(co, H->eta, H->h, T, ap);
  float sum = 0.f;
  for (int i = 0; i <= HAIR_PMAX; ++i) sum += (ap[i].x + ap[i].y + ap[i].z) / 3.f;
  for (int i = 0; i <= HAIR_PMAX; ++i) out[i] = ((ap[i].x + ap[i].y + ap[i].z) / 3.f) / sum;
}
static void hair_rot(const hair_t *H, int p, float so, float co, float *sp, float *cp) {
  if (p == 0) {
    *sp = so * H->cos2k[1] - co * H->sin2k[1];
    *cp = co * H->cos2k[1] + so * H->sin2k[1];
  } else if (p == 1) {
    *sp = so * H->cos2k[0] + co * H->sin2k[0];
    *cp = co * H->cos2k[0] - so * H->sin2k[0];
  } else if (p == 2) {
    *sp = so * H->cos2k[2] + co * H->sin2k[2];
    *cp = co * H->cos2k[2] - so * H->sin2k[2];
  } else {
    *sp = so;
    *cp = co;
  }
  *cp = fabsf(*cp);
}
/* HairMarschner::evalMarschner (hairMarsch.cpp:30-104), local frame of the normal */
static v3 hair_eval(const hair_t *H, v3 wi, v3 wo) {
  float so = wo.x, co = safe_sqrt(1 - so * so);
  float phio = atan2f(wo.z, wo.y);
  float go = asinf(clampf(so, -1.f, 1.f));
  float si = wi.x, ci = safe_sqrt(1 - si * si);
  float phii = atan2f(wi.z, wi.y);
  float st = so / H->eta, ct = safe_sqrt(1 - st * st);
  float etap = safe_sqrt(sqr(H->eta) - sqr(so));
  float sgt = H->h / etap;
  float cgt = safe_sqrt(1 - sgt * sgt);
  float gt = asinf(clampf(sgt, -1.f, 1.f));
  v3 T = V3(fast_exp(-H->sigma_a.x * (2 * cgt / ct)), fast_exp(-H->sigma_a.y * (2 * cgt / ct)),
            fast_exp(-H->sigma_a.z * (2 * cgt / ct)));
  float phi = phii - phio;
  v3 ap[HAIR_PMAX + 1];
  hair_Ap(co, H->eta, H->h, T, ap);
  v3 bsdf = vs(0.f);
  for (int p = 0; p < HAIR_PMAX; ++p) {
    float mp = hair_Mp(ci, co, si, so, H->v[p]);
    float np = hair_Np(phi, p, H->s, go, gt);
    bsdf = vadd(bsdf, vscale(ap[p], mp * np));
  }
  bsdf = vadd(bsdf, vdivs(vscale(ap[HAIR_PMAX], hair_Mp(ci, co, si, so, H->v[HAIR_PMAX])), (float)(2 * M_PI)));
  if (fabsf(wi.z) > 0.f) bsdf = vdivs(bsdf, fabsf(wi.z));
  return bsdf;
}
v3 or_hair_f(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  hair_t H = hair_params(m);
  onb_t b = onb_make(h->n);
  return hair_eval(&H, onb_to_local(&b, vnorm(wi)), onb_to_local(&b, vnorm(wo)));
}
/* HairPDF::at (hairPDF.cpp:83-153) */
static float hair_pdf_local(const hair_t *H, v3 wo_l, v3 wi_l) {
  float si = wi_l.x, ci = safe_sqrt(1 - si * si);
  float phii = atan2f(wi_l.z, wi_l.y);
  float so = wo_l.x, co = safe_sqrt(1 - so * so);
  float phio = atan2f(wo_l.z, wo_l.y);
  float go = asinf(clampf(so, -1.f, 1.f));
  float etap = safe_sqrt(sqr(H->eta) - sqr(so)) / co;
  float sgt = H->h / etap;
  float gt = asinf(clampf(sgt, -1.f, 1.f));
  float dphi = phii - phio;
  if (dphi < -M_PI) dphi = (float)(dphi + 2 * M_PI);
  else if (dphi > M_PI) dphi = (float)(dphi - 2 * M_PI);
  float apdf[HAIR_PMAX + 1];
  hair_ap_pdf(H, co, apdf);
  float pdf = 0.f;
  for (int p = 0; p < HAIR_PMAX; ++p) {
    float sp, cp;
    hair_rot(H, p, so, co, &sp, &cp);
    pdf += apdf[p] * hair_Mp(ci, cp, si, sp, H->v[p]) * hair_Np(dphi, p, H->s, go, gt);
  }
  return pdf;
}
float or_hair_pdf(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  hair_t H = hair_params(m);
  onb_t b = onb_make(h->n);
  return hair_pdf_local(&H, onb_to_local(&b, wo), onb_to_local(&b, vnorm(wi)));
}
/* HairMarschner::scatter + HairPDF::get (hairMarsch.cpp:4-28, hairPDF.cpp:9-81);
 * the tangent frame falls back to the normal (hairMarsch.cpp:13-16). */
int or_hair_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float uc, float u0,
                    float u1, or_srec *s) {
  hair_t H = hair_params(m);
  v3 wo = vneg(vnorm(ray_d));
  /* HairMarschner::scatter (hairMarsch.cpp:4-28): the sample and pdf_val frame
   * on hair_tangent, the normal when it is unset (near_zero); attenuation is
   * f(), which frames on the normal (hairMarsch.hpp:44-54) */
  const int has_t = h->tangent.x != 0.f || h->tangent.y != 0.f || h->tangent.z != 0.f;
  onb_t b = onb_make(has_t ? h->tangent : h->n);
  v3 wo_l = onb_to_local(&b, wo);
  float so = wo_l.x, co = safe_sqrt(1 - so * so);
  float phio = atan2f(wo_l.z, wo_l.y);
  float go = asinf(clampf(so, -1.f, 1.f));
  float apdf[HAIR_PMAX + 1];
  hair_ap_pdf(&H, co, apdf);
  int p = sample_discrete(apdf, HAIR_PMAX + 1, uc, NULL, NULL);
  float sp, cp;
  hair_rot(&H, p, so, co, &sp, &cp);
  float v = H.v[p];
  float cth = 1 + v * logf(fmaxf(u0, (float)1e-5) + (1.f - u0) * fast_exp(-2.f / v));
  float sth = safe_sqrt(1 - sqr(cth));
  float cphi = (float)cos(2 * M_PI * u1);
  float si = -cth * sp + sth * cphi * cp;
  float ci = safe_sqrt(1 - sqr(si));
  float etap = safe_sqrt(sqr(H.eta) - sqr(so)) / co;
  float sgt = H.h / etap;
  float gt = asinf(clampf(sgt, -1.f, 1.f));
  float dphi;
  if (p < HAIR_PMAX)
    dphi = hair_phi(p, go, gt) + sample_trimmed_logistic(uc, H.s, (float)-M_PI, (float)M_PI);
  else
    dphi = (float)(2 * M_PI * uc);
  float phii = dphi + phio;
  v3 wi_l = V3(si, ci * cosf(phii), ci * sinf(phii));
  v3 wi = vnorm(onb_to_world(&b, wi_l));
  if (has_t) {
    onb_t bn = onb_make(h->n);
    s->att = hair_eval(&H, onb_to_local(&bn, wi), onb_to_local(&bn, wo));
  } else {
    s->att = hair_eval(&H, onb_to_local(&b, wi), onb_to_local(&b, wo));
  }
  s->pdf = hair_pdf_local(&H, wo_l, onb_to_local(&b, wi));
  s->wi = wi;
  s->flags = OR_GLOSSY | OR_REFL;
  return 1;
}
