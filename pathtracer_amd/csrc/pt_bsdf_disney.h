// pt_bsdf_disney.h -- DisneyPrincipledBSDF and HairMarschner on the device.
//
//   Disney mixture       disneyPrincipledBSDF.hpp:8-99, disneyPrincipledBSDF.cpp:3-114
//   lobes                disneyDiffuse.hpp:8-88, disneyMetal.hpp:9-140, disneyGlass.hpp:5-168,
//                        disneySheen.hpp:7-72, disneyClearcoat.hpp:7-79, clearcoatPDF.hpp:6-61,
//                        ClearcoatTRD materialMath.hpp:216-279
//   hair                 hairMarsch.hpp:3-158, hairMarsch.cpp:4-117, hairPDF.cpp:4-153
//
// Reference behaviour kept: front faces (is_inside) use the glass lobe only;
// the mixture's metal lobe is f_variant with default arguments (C0 =
// 0.5*baseColor); the chosen lobe samples with u = (uc_remapped, u.y);
// specular lobes skip the pmf; sheen is absent from the pdf; the hair's
// evalMarschner uses the unrotated theta_o in Mp and its pdf sums p < pMax.
#pragma once
#include "pt_math.h"

namespace pt {

PT_D float rgb_luminance(f3 c) { return dot3(c, mk3(0.2126f, 0.7152f, 0.0722f)); }
// x^N for a constant integer N by repeated squaring in double.  Rounded to
// float, the product (relative error below N 2^-53) is the float nearest x^N,
// the value a correctly rounded pow gives, unless x^N lies within that error
// of the midpoint between two floats.  A fraction of the instructions of
// powf, which the Schlick weights call several times per Disney evaluation.
template <int N>
PT_D double powi_d(double x) {
  if constexpr (N == 1) {
    return x;
  } else if constexpr (N % 2 == 0) {
    const double h = powi_d<N / 2>(x);
    return h * h;
  } else {
    return powi_d<N - 1>(x) * x;
  }
}
PT_D float pow5_rn(double x) { return (float)powi_d<5>(x); }
// materialMath.hpp:55-63: fresnelSchlick raises the double 1.0 - cos to the
// fifth power, hackedSchlick the float 1.f - cos
PT_D f3 fresnel_schlick(float c, f3 F0) {
  float w = pow5_rn(1.0 - (double)c);
  return mk3(F0.x + (1.f - F0.x) * w, F0.y + (1.f - F0.y) * w, F0.z + (1.f - F0.z) * w);
}
PT_D float hacked_schlick(float c, float F90) { return 1.f + (F90 - 1.f) * pow5_rn((double)(1.f - c)); }
PT_D TRD trd_from_roughness(float roughness, float aniso) {
  aniso = fmaxf(0.0f, fminf(aniso, 1.0f / 0.9f - 0.00001f));
  float aspect = safe_sqrtf(1.0f - 0.9f * aniso);
  if (aspect == 0.0f) aspect = 0.0001f;
  TRD d;
  d.ax = fmaxf(0.0001f, sqrf(roughness) / aspect);
  d.ay = fmaxf(0.0001f, sqrf(roughness) * aspect);
  return d;
}

struct Disney {
  f3 base;
  float spec_trans, metallic, subsurface, roughness, sheen_tint, eta;
  float dw, mw, gw, cw, sw;
  TRD dist;
  float cc_alpha;
};
PT_D Disney disney_params(const pt_material_desc &m, const ShadePoint &h) {
  Disney d;
  d.base = h.has_tex ? h.tex : mk3(m.p[0], m.p[1], m.p[2]);
  d.spec_trans = m.p[3];
  d.metallic = m.p[4];
  d.subsurface = m.p[5];
  d.roughness = m.p[7];
  float anisotropic = m.p[9], sheen = m.p[10], clearcoat = m.p[12], gloss = m.p[13];
  d.sheen_tint = m.p[11];
  d.eta = m.p[14];
  d.dw = (1 - d.metallic) * (1 - d.spec_trans);
  d.mw = (1 - d.spec_trans * (1 - d.metallic));
  d.gw = (1 - d.metallic) * d.spec_trans;
  d.cw = 0.25f * clearcoat;
  d.sw = (1 - d.metallic) * sheen;
  d.dist = trd_from_roughness(d.roughness, anisotropic);
  d.cc_alpha = (1.f - gloss) * 0.1f + gloss * 0.001f;
  return d;
}

PT_D f3 ddiffuse_f(const Disney &d, f3 wo_w, f3 wi_w, const ShadePoint &h) {
  Frame b = make_frame(h.n);
  f3 wo = to_local(b, wo_w), wi = to_local(b, wi_w);
  float aci = fabsf(wi.z), aco = fabsf(wo.z);
  f3 wh = norm3(add3(wo, wi));
  float whwi = dot3(wh, wi);
  float fd90 = 0.5f + 2.f * d.roughness * powf(fabsf(whwi), 2.f);
  f3 bd = scl3(scl3(d.base, 1.f / PT_PI), hacked_schlick(aci, fd90) * hacked_schlick(aco, fd90));
  float fss90 = d.roughness * powf(fabsf(whwi), 2.f);
  f3 bs = scl3(scl3(d.base, 1.25f), 1.f / PT_PI);
  bs = scl3(bs, hacked_schlick(aci, fss90) * hacked_schlick(aco, fss90) * (1.f / (aci + aco) - 0.5f) + 0.5f);
  return add3(scl3(bd, 1.f - d.subsurface), scl3(bs, d.subsurface));
}
PT_D f3 dmetal_f_c0(const Disney &d, f3 wo_w, f3 wi_w, const ShadePoint &h, f3 c0) {
  Frame b = make_frame(h.n);
  f3 wo = to_local(b, wo_w), wi = to_local(b, wi_w);
  f3 wh = norm3(add3(wo, wi));
  if (wi.z <= 0.f || wo.z <= 0.f || trd_smooth(d.dist)) return s3(0.f);
  f3 F = fresnel_schlick(fabsf(dot3(wh, wi)), c0);
  float D = trd_D(d.dist, wh), G = trd_G(d.dist, wo, wi);
  return divs3(scl3(scl3(F, D), G), fabsf(4.f * wi.z * wo.z));
}
PT_D f3 dmetal_f_variant(const Disney &d, f3 wo, f3 wi, const ShadePoint &h) {
  // f_variant(specular_tint .5, eta 1, specular .5, metallic .5): R0 = 0, C0 = 0.5 base
  f3 ks = add3(s3(0.5f), scl3(d.base, 0.5f));
  f3 c0 = add3(scl3(scl3(ks, 0.f * (1.f - 0.5f)), 0.5f), scl3(d.base, 0.5f));
  return dmetal_f_c0(d, wo, wi, h, c0);
}
PT_D float dmetal_pdf(const Disney &d, f3 wo, f3 wi, const ShadePoint &h) {
  if (trd_smooth(d.dist)) return 0.f;
  Frame b = make_frame(h.n);
  return mf_pdf_at(d.dist, b, to_local(b, norm3(wo)), wi);
}
// ATT: the rough lobes' value at the sampled direction.  disney_scatter
// re-evaluates every lobe there (disney_f_pdf) and keeps only the sampled
// lobe's pdf for its cut-off, so it asks for the value only where the lobe is
// specular.
template <bool ATT>
PT_D bool dmetal_scatter(const Disney &d, const Frame &b, f3 wo_l, f3 wo, const ShadePoint &h, float u0, float u1,
                         Scatter &s) {
  if (trd_smooth(d.dist)) {
    f3 wi = norm3(reflect_pbrt(wo, h.n));
    if (dot3(h.n, wi) <= 0.f) return false;
    s.wi = wi;
    s.att = fresnel_schlick(fabsf(dot3(h.n, wi)), d.base);
    s.flags = BX_SPECULAR | BX_REFL;
    s.pdf = 0.f;
    return true;
  }
  f3 ml = trd_sample_wm(d.dist, wo_l, u0, u1);
  f3 wi_l = reflect_pbrt(wo_l, ml);
  if (!same_hemi(wo_l, wi_l)) wi_l.z = -wo_l.z;
  f3 wi = norm3(to_world(b, wi_l));
  if (near_zero3(wi)) return false;
  if (dot3(h.n, wi) <= 0.f) return false;
  s.wi = norm3(wi);
  s.flags = BX_GLOSSY | BX_REFL;
  s.att = ATT ? dmetal_f_c0(d, wo, wi, h, d.base) : s3(0.f);
  s.pdf = mf_pdf_at(d.dist, b, wo_l, wi);
  return true;
}
PT_D f3 dglass_f(const Disney &d, f3 wo, f3 wi, const ShadePoint &h) {
  float eta = d.eta;
  TRD t = d.dist;
  if (eta == 1.f || trd_smooth(t)) return s3(0.f);
  Frame b = make_frame(h.n);
  f3 wo_l = to_local(b, norm3(wo)), wi_l = to_local(b, norm3(wi));
  float co = wo_l.z, ci = wi_l.z;
  bool refl = ci * co > 0.f;
  float etap = 1.f;
  if (!refl) etap = co > 0.f ? eta : (1.f / eta);
  f3 wm = add3(scl3(wi_l, etap), wo_l);
  if (ci == 0.f || co == 0.f || len3(wm) == 0.f) return s3(0.f);
  wm = face_fwd(norm3(wm), mk3(0.f, 0.f, 1.f));
  if (dot3(wm, wi_l) * ci < 0.f || dot3(wm, wo_l) * co < 0.f) return s3(0.f);
  float F = fresnel_diel(dot3(wo_l, wm), eta);
  if (refl) return scl3(d.base, trd_D(t, wm) * trd_G(t, wo_l, wi_l) * F / fabsf(4.f * ci * co));
  float den = sqrf(dot3(wi_l, wm) + dot3(wo_l, wm) / etap) * ci * co;
  f3 sc = mk3(safe_sqrtf(d.base.x), safe_sqrtf(d.base.y), safe_sqrtf(d.base.z));
  return scl3(sc, trd_D(t, wm) * (1.f - F) * trd_G(t, wo_l, wi_l) * fabsf(dot3(wi_l, wm) * dot3(wo_l, wm) / den));
}
PT_D float dglass_pdf(const Disney &d, f3 wo, f3 wi, const ShadePoint &h) {
  if (d.eta == 1.f || trd_smooth(d.dist)) return 0.f;
  Frame b = make_frame(h.n);
  return mfdiel_pdf_local(d.eta, d.dist, to_local(b, norm3(wo)), to_local(b, norm3(wi)));
}
template <bool ATT>
PT_D bool dglass_scatter(const Disney &d, const Frame &b, f3 wo_l, f3 wo, const ShadePoint &h, float uc, float u0,
                         float u1, Scatter &s) {
  float eta = d.eta;
  TRD t = d.dist;
  if (eta == 1.f || trd_smooth(t)) {
    float co = dot3(h.n, wo);
    float R = fresnel_diel(co, eta), T = 1.f - R;
    if (R == 0.f && T == 0.f) return false;
    if (uc < R / (R + T)) {
      s.wi = norm3(reflect_pbrt(wo, h.n));
      s.flags = BX_SPECULAR | BX_REFL;
      s.att = s3(1.f);
      s.pdf = R;
      return true;
    }
    float etap;
    f3 wi;
    if (!refract_pbrt(wo, h.n, eta, &etap, &wi)) return false;
    s.wi = wi;
    s.flags = BX_SPECULAR | BX_TRANS;
    s.att = s3(1.f);
    s.pdf = T;
    return true;
  }
  f3 ml = trd_sample_wm(t, wo_l, u0, u1);
  float R = fresnel_diel(dot3(wo_l, ml), eta), T = 1.f - R;
  if (R + T == 0.f) return false;
  f3 wi_l;
  if (uc < R / (R + T)) {
    wi_l = reflect_pbrt(wo_l, ml);
    if (!same_hemi(wo_l, wi_l)) wi_l.z = -wi_l.z;
  } else {
    float es;
    bool valid = refract_pbrt(wo_l, ml, eta, &es, &wi_l);
    if (!valid || same_hemi(wo_l, wi_l)) wi_l = reflect_pbrt(wo_l, ml);
  }
  f3 wi = to_world(b, wi_l);
  if (near_zero3(wi)) return false;
  bool is_refl = dot3(to_local(b, wi), to_local(b, wo)) > 0.f;
  s.wi = wi;
  s.flags = BX_GLOSSY | (is_refl ? BX_REFL : BX_TRANS);
  s.att = ATT ? dglass_f(d, wo, wi, h) : s3(0.f);
  s.pdf = mfdiel_pdf_local(eta, t, wo_l, to_local(b, norm3(wi)));
  return true;
}
PT_D float cc_D(float a, f3 hh) {
  float upper = a * a - 1;
  float lower = PT_PI * logf(sqrf(a));
  float term = 1 + (sqrf(a) - 1) * sqrf(hh.z);
  return upper / (lower * term);
}
PT_D float cc_lambda(f3 w) {
  float t2 = tan2t(w);
  if (isinf(t2)) return 0.f;
  float ca = cosphi(w) * 0.25f, sa = sinphi(w) * 0.25f;
  float a2 = ca * ca + sa * sa;
  return (sqrtf(1 + a2 * t2) - 1) / 2;
}
PT_D f3 dcc_f(const Disney &d, f3 wo_w, f3 wi_w, const ShadePoint &h) {
  Frame b = make_frame(h.n);
  f3 wo = to_local(b, wo_w), wi = to_local(b, wi_w);
  f3 wh = norm3(add3(wo, wi));
  if (wi.z <= 0.f || wo.z <= 0.f) return s3(0.f);
  float r0 = sqrf((1.5f - 1) / (1.5f + 1));
  f3 Fc = fresnel_schlick(fabsf(dot3(wh, wi)), s3(r0));
  float Dc = cc_D(d.cc_alpha, wh), Gc = 1.f / (1.f + cc_lambda(wo) + cc_lambda(wi));
  return divs3(scl3(scl3(Fc, Dc), Gc), fabsf(4.f * wi.z * wo.z));
}
PT_D float cc_pdf_at(float a, const Frame &b, f3 wo_l, f3 wi_w) {
  f3 wi = to_local(b, norm3(wi_w));
  if (wi.z <= 0.f) return 0.f;
  f3 hh = norm3(add3(wo_l, wi));
  float den = 4.f * fabsf(dot3(wi, hh));
  if (den < 1e-6f) return 0.f;
  return cc_D(a, hh) * fabsf(hh.z) / den;
}
PT_D float dcc_pdf(const Disney &d, f3 wo, f3 wi, const ShadePoint &h) {
  Frame b = make_frame(h.n);
  return cc_pdf_at(d.cc_alpha, b, to_local(b, norm3(wo)), wi);
}
template <bool ATT>
PT_D bool dcc_scatter(const Disney &d, const Frame &b, f3 wo_l, f3 wo, const ShadePoint &h, float u0, float u1,
                      Scatter &s) {
  float a = d.cc_alpha;
  float upper = 1 - powf(sqrf(a), 1 - u0);
  float lower = 1 - sqrf(a);
  float ch = safe_sqrtf(upper / lower);
  float sh = safe_sqrtf(1 - sqrf(ch));
  float az = 2.f * PT_PI * u1;
  f3 hl = mk3(sh * cosf(az), sh * sinf(az), ch);
  f3 wi_l = reflect_pbrt(wo_l, hl);
  if (!same_hemi(wo_l, wi_l)) wi_l.z = -wo_l.z;  // BIASED_SAMPLING (clearcoatPDF.hpp:27-29)
  f3 wi = norm3(to_world(b, wi_l));
  if (dot3(h.n, wi) <= 0.f) return false;
  s.wi = norm3(wi);
  s.flags = BX_GLOSSY | BX_REFL;
  s.att = ATT ? dcc_f(d, wo, wi, h) : s3(0.f);
  s.pdf = cc_pdf_at(a, b, wo_l, wi);
  return true;
}
PT_D f3 dsheen_f(const Disney &d, f3 wo_w, f3 wi_w, const ShadePoint &h) {
  Frame b = make_frame(h.n);
  f3 wo = to_local(b, wo_w), wi = to_local(b, wi_w);
  f3 wh = norm3(add3(wo, wi));
  float a = fabsf(dot3(wh, wi));
  float lum = rgb_luminance(d.base);
  f3 ct = lum > 0.f ? divs3(d.base, lum) : s3(1.f);
  f3 cs = add3(s3(1.f), scl3(sub3(ct, s3(1.f)), d.sheen_tint));
  return scl3(cs, pow5_rn((double)(1 - a)));
}

PT_D f3 disney_f(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) {
  Disney d = disney_params(m, h);
  if (h.inside) return dglass_f(d, wo, wi, h);
  f3 r = scl3(ddiffuse_f(d, wo, wi, h), d.dw);
  r = add3(r, scl3(dmetal_f_variant(d, wo, wi, h), d.mw));
  r = add3(r, scl3(dglass_f(d, wo, wi, h), d.gw));
  r = add3(r, scl3(dcc_f(d, wo, wi, h), d.cw));
  r = add3(r, scl3(dsheen_f(d, wo, wi, h), d.sw));
  return r;
}
PT_D float disney_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) {
  Disney d = disney_params(m, h);
  if (h.inside) return dglass_pdf(d, wo, wi, h);
  float sum = d.dw + d.mw + d.gw + d.cw;
  float p = (d.dw / sum) * cosine_pdf(h.n, wi);
  p += (d.mw / sum) * dmetal_pdf(d, wo, wi, h);
  p += (d.gw / sum) * dglass_pdf(d, wo, wi, h);
  p += (d.cw / sum) * dcc_pdf(d, wo, wi, h);
  return p;
}
// Glass lobe f and pdf from the normalised local directions (dglass_f and
// dglass_pdf's operations, shared terms once).
PT_D void dglass_f_pdf_local(const Disney &d, f3 wo_l, f3 wi_l, float lo, f3 &f, float &pdf) {
  f = s3(0.f);
  pdf = 0.f;
  const float eta = d.eta;
  const TRD t = d.dist;
  if (eta == 1.f || trd_smooth(t)) return;
  float co = wo_l.z, ci = wi_l.z;
  bool refl = ci * co > 0.f;
  float etap = 1.f;
  if (!refl) etap = co > 0.f ? eta : (1.f / eta);
  f3 wm = add3(scl3(wi_l, etap), wo_l);
  if (ci == 0.f || co == 0.f || len3(wm) == 0.f) return;
  wm = face_fwd(norm3(wm), mk3(0.f, 0.f, 1.f));
  if (dot3(wm, wi_l) * ci < 0.f || dot3(wm, wo_l) * co < 0.f) return;
  const float F = fresnel_diel(dot3(wo_l, wm), eta), T = 1.f - F;
  const float D = trd_D(t, wm);  // lo = trd_lambda(t, wo_l)
  const float G = 1.f / (1.f + lo + trd_lambda(t, wi_l));
  const float Dw = 1.f / (1.f + lo) / abs_cost(wo_l) * D * fabsf(dot3(wo_l, wm));
  if (refl) {
    f = scl3(d.base, D * G * F / fabsf(4.f * ci * co));
    if (F + T != 0.f) pdf = Dw / (4.f * fabsf(dot3(wo_l, wm))) * F / (F + T);
    return;
  }
  const float den = sqrf(dot3(wi_l, wm) + dot3(wo_l, wm) / etap) * ci * co;
  f3 sc = mk3(safe_sqrtf(d.base.x), safe_sqrtf(d.base.y), safe_sqrtf(d.base.z));
  f = scl3(sc, D * (1.f - F) * G * fabsf(dot3(wi_l, wm) * dot3(wo_l, wm) / den));
  if (F + T != 0.f) {
    const float den_p = sqrf(dot3(wi_l, wm) + dot3(wo_l, wm) / etap);
    pdf = Dw * (fabsf(dot3(wi_l, wm)) / den_p) * T / (F + T);
  }
}

// What the scatter and every evaluation of one shading point share: the
// parameters, the shading frame, wo in it (as given and normalised, as the
// lobes use them) and wo's masking terms.
struct DisneyCtx {
  Disney d;
  Frame b;
  f3 wo_n, wo_r;
  float lam_n, lam_r, cc_lam_r;  // trd_lambda(wo_n), trd_lambda(wo_r), cc_lambda(wo_r)
};
PT_D DisneyCtx disney_ctx(const pt_material_desc &m, const ShadePoint &h, f3 wo) {
  DisneyCtx c;
  c.d = disney_params(m, h);
  c.b = make_frame(h.n);
  c.wo_n = to_local(c.b, norm3(wo));
  c.wo_r = to_local(c.b, wo);
  c.lam_n = trd_lambda(c.d.dist, c.wo_n);
  c.lam_r = trd_lambda(c.d.dist, c.wo_r);
  c.cc_lam_r = cc_lambda(c.wo_r);
  return c;
}
// disney_f and disney_pdf in one pass: the half vectors and the shared
// microfacet terms once, wo's terms from the context.  Every lobe keeps its
// own early-outs and operation order, so f and pdf equal the separate
// functions' bit for bit.
PT_D void disney_f_pdf_ctx(const DisneyCtx &c, f3 wi, const ShadePoint &h, f3 &f, float &pdf) {
  const Disney &d = c.d;
  const Frame &b = c.b;
  const f3 wo_n = c.wo_n, wi_n = to_local(b, norm3(wi));
  f3 gf;
  float gp;
  dglass_f_pdf_local(d, wo_n, wi_n, c.lam_n, gf, gp);
  if (h.inside) {
    f = gf;
    pdf = gp;
    return;
  }
  const f3 wo_r = c.wo_r, wi_r = to_local(b, wi);  // ddiffuse_f, dmetal_f_c0, dcc_f, dsheen_f
  const f3 wh = norm3(add3(wo_r, wi_r));
  const float whwi = dot3(wh, wi_r);
  // diffuse (ddiffuse_f)
  f3 fd;
  {
    float aci = fabsf(wi_r.z), aco = fabsf(wo_r.z);
    float fd90 = 0.5f + 2.f * d.roughness * powf(fabsf(whwi), 2.f);
    f3 bd = scl3(scl3(d.base, 1.f / PT_PI), hacked_schlick(aci, fd90) * hacked_schlick(aco, fd90));
    float fss90 = d.roughness * powf(fabsf(whwi), 2.f);
    f3 bs = scl3(scl3(d.base, 1.25f), 1.f / PT_PI);
    bs = scl3(bs, hacked_schlick(aci, fss90) * hacked_schlick(aco, fss90) * (1.f / (aci + aco) - 0.5f) + 0.5f);
    fd = add3(scl3(bd, 1.f - d.subsurface), scl3(bs, d.subsurface));
  }
  // metal (dmetal_f_variant -> dmetal_f_c0; trd_G with wo's lambda)
  f3 fm = s3(0.f);
  if (!(wi_r.z <= 0.f || wo_r.z <= 0.f || trd_smooth(d.dist))) {
    f3 ks = add3(s3(0.5f), scl3(d.base, 0.5f));
    f3 c0 = add3(scl3(scl3(ks, 0.f * (1.f - 0.5f)), 0.5f), scl3(d.base, 0.5f));
    f3 F = fresnel_schlick(fabsf(whwi), c0);
    float D = trd_D(d.dist, wh), G = 1.f / (1.f + c.lam_r + trd_lambda(d.dist, wi_r));
    fm = divs3(scl3(scl3(F, D), G), fabsf(4.f * wi_r.z * wo_r.z));
  }
  // clearcoat (dcc_f)
  f3 fc = s3(0.f);
  if (!(wi_r.z <= 0.f || wo_r.z <= 0.f)) {
    float r0 = sqrf((1.5f - 1) / (1.5f + 1));
    f3 Fc = fresnel_schlick(fabsf(whwi), s3(r0));
    float Dc = cc_D(d.cc_alpha, wh), Gc = 1.f / (1.f + c.cc_lam_r + cc_lambda(wi_r));
    fc = divs3(scl3(scl3(Fc, Dc), Gc), fabsf(4.f * wi_r.z * wo_r.z));
  }
  // sheen (dsheen_f)
  f3 fs;
  {
    float a = fabsf(whwi);
    float lum = rgb_luminance(d.base);
    f3 ct = lum > 0.f ? divs3(d.base, lum) : s3(1.f);
    f3 cs = add3(s3(1.f), scl3(sub3(ct, s3(1.f)), d.sheen_tint));
    fs = scl3(cs, pow5_rn((double)(1 - a)));
  }
  f3 r = scl3(fd, d.dw);
  r = add3(r, scl3(fm, d.mw));
  r = add3(r, scl3(gf, d.gw));
  r = add3(r, scl3(fc, d.cw));
  r = add3(r, scl3(fs, d.sw));
  f = r;
  // pdf: the lobes' pdfs at the normalised directions
  float pm = 0.f, pc = 0.f;
  if (wi_n.z > 0.f) {
    const f3 hn = norm3(add3(wo_n, wi_n));
    if (!trd_smooth(d.dist)) {  // dmetal_pdf -> mf_pdf_at -> trd_Dw with wo's G1
      float den = 4.f * fabsf(dot3(wo_n, hn));
      if (!(den < 1e-6f))
        pm = 1.f / (1.f + c.lam_n) / abs_cost(wo_n) * trd_D(d.dist, hn) * fabsf(dot3(wo_n, hn)) / den;
    }
    {  // dcc_pdf -> cc_pdf_at
      float den = 4.f * fabsf(dot3(wi_n, hn));
      if (!(den < 1e-6f)) pc = cc_D(d.cc_alpha, hn) * fabsf(hn.z) / den;
    }
  }
  float sum = d.dw + d.mw + d.gw + d.cw;
  float p = (d.dw / sum) * cosine_pdf(h.n, wi);
  p += (d.mw / sum) * pm;
  p += (d.gw / sum) * gp;
  p += (d.cw / sum) * pc;
  pdf = p;
}
PT_D void disney_f_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h, f3 &f, float &pdf) {
  disney_f_pdf_ctx(disney_ctx(m, h, wo), wi, h, f, pdf);
}

// gl::sampleDiscrete (utility.hpp:155-190)
template <int N>
PT_D int sample_discrete(const float (&w)[N], float u, float *pmf, float *u_remap) {
  float sum = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) sum += w[i];
  float up = u * sum;
  if (up == sum) up = nextafterf(up, -INFINITY);
  int off = 0;
  float acc = 0.f;
  while (off < N - 1 && acc + w[off] <= up) {
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
PT_D bool disney_scatter_ctx(const DisneyCtx &c, f3 wo, const ShadePoint &h, float uc, float u0, float u1,
                             Scatter &s) {
  const Disney &d = c.d;
  if (h.inside) return dglass_scatter<true>(d, c.b, c.wo_n, wo, h, uc, u0, u1, s);
  float w[4] = {d.dw, d.mw, d.gw, d.cw};
  float pmf, ucr;
  int lobe = sample_discrete<4>(w, uc, &pmf, &ucr);
  if (pmf <= 1e-6f) return false;
  bool ok;
  if (lobe == 0) {
    f3 wiw = to_world(c.b, cosine_hemi(ucr, u1));
    s.wi = norm3(wiw);
    s.pdf = cosine_pdf(h.n, wiw);
    s.flags = BX_DIFFUSE | BX_REFL;
    s.att = s3(0.f);  // re-evaluated below (ddiffuse_f's value is not kept)
    ok = true;
  } else if (lobe == 1) {
    ok = dmetal_scatter<false>(d, c.b, c.wo_n, wo, h, ucr, u1, s);
  } else if (lobe == 2) {
    ok = dglass_scatter<false>(d, c.b, c.wo_n, wo, h, ucr, ucr, u1, s);
  } else {
    ok = dcc_scatter<false>(d, c.b, c.wo_n, wo, h, ucr, u1, s);
  }
  if (!ok) return false;
  if (s.flags & BX_SPECULAR) return true;
  if (s.pdf <= 1e-6f) return false;
  f3 wiw = norm3(s.wi);
  disney_f_pdf_ctx(c, wiw, h, s.att, s.pdf);
  return s.pdf > 1e-7f;
}
PT_D bool disney_scatter(const pt_material_desc &m, f3 wo, const ShadePoint &h, float uc, float u0, float u1,
                         Scatter &s) {
  return disney_scatter_ctx(disney_ctx(m, h, wo), wo, h, uc, u0, u1, s);
}

// ============================== HairMarschner
#define HAIR_PMAX 3
struct Hair {
  f3 sigma_a;
  float h, eta;
  float v[HAIR_PMAX + 1];
  float s;
  float sin2k[HAIR_PMAX], cos2k[HAIR_PMAX];
};
PT_D Hair hair_params(const pt_material_desc &m) {
  Hair H;
  H.sigma_a = mk3(m.p[0], m.p[1], m.p[2]);
  H.h = m.p[3];
  H.eta = m.p[4];
  float bm = m.p[5], bn = m.p[6], alpha = m.p[7];
  H.v[0] = sqrf(0.726f * bm + 0.812f * sqrf(bm) + 3.7f * (float)powi_d<20>((double)bm));
  H.v[1] = 0.25f * H.v[0];
  H.v[2] = 4 * H.v[0];
  H.v[3] = H.v[2];
  H.s = 0.626657069f * (0.265f * bn + 1.194f * sqrf(bn) + 5.372f * (float)powi_d<22>((double)bn));
  H.sin2k[0] = sinf(alpha);
  H.cos2k[0] = safe_sqrtf(1 - sqrf(H.sin2k[0]));
#pragma unroll
  for (int i = 1; i < HAIR_PMAX; ++i) {
    H.sin2k[i] = 2 * H.cos2k[i - 1] * H.sin2k[i - 1];
    H.cos2k[i] = sqrf(H.cos2k[i - 1]) - sqrf(H.sin2k[i - 1]);
  }
  return H;
}
PT_D float i0f(float x) {
  float term = 1.0f, sum = term, x2 = x * x;
  for (int k = 1; k < 500; ++k) {
    term *= x2 / (4.0f * k * k);
    sum += term;
    if (term < 1e-6f * sum) break;
    if (sum > 3.402823466e+38f * 0.5f) break;
  }
  return sum;
}
PT_D float logi0f(float x) {
  if (x > 12.0f) return x - 0.5f * logf(2.0f * PT_PI * x);
  return logf(i0f(x));
}
// gl::fastExp (utility.hpp:248-268); saturates where the reference throws
PT_D float fast_exp(float x) {
  float xp = x * 1.442695041f;
  float fxp = floorf(xp);
  float f = xp - fxp;
  int i = (int)fxp;
  float t = fmaf(f, fmaf(f, fmaf(f, 0.0781455737f, 0.226173572f), 0.695556856f), 1.f);
  uint32_t bits = __float_as_uint(t);
  int e = (int)(bits >> 23) - 127 + i;
  if (e < -126) return 0.f;
  if (e > 127) return __int_as_float(0x7f800000);
  bits &= 0x807FFFFFu;
  bits |= (uint32_t)(e + 127) << 23;
  return __uint_as_float(bits);
}
PT_D float trimmed_logistic(float x, float s, float a, float b) {
  x = fabsf(x);
  float lg = expf(-x / s) / (s * sqrf(1 + expf(-x / s)));
  return lg / (1 / (1 + expf(-b / s)) - 1 / (1 + expf(-a / s)));
}
PT_D float sample_trimmed_logistic(float u, float s, float a, float b) {
  float pa = 1 / (1 + expf(-a / s)), pb = 1 / (1 + expf(-b / s));
  u = pa + u * (pb - pa);
  float x = -s * logf(1 / u - 1);
  return clampf_(x, a, b);
}
PT_D float hair_Mp(float ci, float co, float si, float so, float v) {
  float a = ci * co / v, b = si * so / v;
  if ((double)v <= .1) return fast_exp(logi0f(a) - b - 1 / v + 0.6931f + logf(1 / (2 * v)));
  return (fast_exp(-b) * i0f(a)) / (sinhf(1 / v) * 2 * v);
}
PT_D float hair_phi(int p, float go, float gt) { return 2 * p * gt - 2 * go + p * PT_PI; }
PT_D float hair_Np(float phi, int p, float s, float go, float gt) {
  float dphi = phi - hair_phi(p, go, gt);
  while (dphi > PT_PI) dphi -= 2.f * PT_PI;
  while (dphi < -PT_PI) dphi += 2.f * PT_PI;
  return trimmed_logistic(dphi, s, -PT_PI, PT_PI);
}
PT_D void hair_Ap(float co, float eta, float h, f3 T, f3 (&ap)[HAIR_PMAX + 1]) {
  float c = co * safe_sqrtf(1 - sqrf(h));
  float f = fresnel_diel(c, eta);
  ap[0] = s3(f);
  ap[1] = scl3(T, sqrf(1 - f));
  ap[2] = scl3(mul3(ap[1], T), f);
  f3 den = sub3(s3(1.f), scl3(T, f));
  if (!near_zero3(den))
    ap[3] = mk3(ap[2].x * f * T.x / den.x, ap[2].y * f * T.y / den.y, ap[2].z * f * T.z / den.z);
  else
    ap[3] = s3(0.f);
}
PT_D void hair_ap_pdf(const Hair &H, float co, float (&out)[HAIR_PMAX + 1]) {
  float so = safe_sqrtf(1 - sqrf(co));
  float st = so / H.eta;
  float ct = safe_sqrtf(1 - sqrf(st));
  float etap = safe_sqrtf(sqrf(H.eta) - sqrf(so)) / co;
  float sgt = H.h / etap;
  float cgt = safe_sqrtf(1 - sqrf(sgt));
  float k = 2 * cgt / ct;
  f3 T = mk3(expf(-H.sigma_a.x * k), expf(-H.sigma_a.y * k), expf(-H.sigma_a.z * k));
  f3 ap[HAIR_PMAX + 1];
  hair_Ap(co, H.eta, H.h, T, ap);
  float sum = 0.f;
#pragma unroll
  for (int i = 0; i <= HAIR_PMAX; ++i) sum += (ap[i].x + ap[i].y + ap[i].z) / 3.f;
#pragma unroll
  for (int i = 0; i <= HAIR_PMAX; ++i) out[i] = ((ap[i].x + ap[i].y + ap[i].z) / 3.f) / sum;
}
PT_D void hair_rot(const Hair &H, int p, float so, float co, float &sp, float &cp) {
  if (p == 0) {
    sp = so * H.cos2k[1] - co * H.sin2k[1];
    cp = co * H.cos2k[1] + so * H.sin2k[1];
  } else if (p == 1) {
    sp = so * H.cos2k[0] + co * H.sin2k[0];
    cp = co * H.cos2k[0] - so * H.sin2k[0];
  } else if (p == 2) {
    sp = so * H.cos2k[2] + co * H.sin2k[2];
    cp = co * H.cos2k[2] - so * H.sin2k[2];
  } else {
    sp = so;
    cp = co;
  }
  cp = fabsf(cp);
}
PT_D f3 hair_eval(const Hair &H, f3 wi, f3 wo) {
  float so = wo.x, co = safe_sqrtf(1 - so * so);
  float phio = atan2f(wo.z, wo.y);
  float go = asinf(clampf_(so, -1.f, 1.f));
  float si = wi.x, ci = safe_sqrtf(1 - si * si);
  float phii = atan2f(wi.z, wi.y);
  float st = so / H.eta, ct = safe_sqrtf(1 - st * st);
  float etap = safe_sqrtf(sqrf(H.eta) - sqrf(so));
  float sgt = H.h / etap;
  float cgt = safe_sqrtf(1 - sgt * sgt);
  float gt = asinf(clampf_(sgt, -1.f, 1.f));
  float k = 2 * cgt / ct;
  f3 T = mk3(fast_exp(-H.sigma_a.x * k), fast_exp(-H.sigma_a.y * k), fast_exp(-H.sigma_a.z * k));
  float phi = phii - phio;
  f3 ap[HAIR_PMAX + 1];
  hair_Ap(co, H.eta, H.h, T, ap);
  f3 bsdf = s3(0.f);
  for (int p = 0; p < HAIR_PMAX; ++p) {
    float mp = hair_Mp(ci, co, si, so, H.v[p]);
    float np = hair_Np(phi, p, H.s, go, gt);
    bsdf = add3(bsdf, scl3(ap[p], mp * np));
  }
  bsdf = add3(bsdf, divs3(scl3(ap[HAIR_PMAX], hair_Mp(ci, co, si, so, H.v[HAIR_PMAX])), 2.f * PT_PI));
  if (fabsf(wi.z) > 0.f) bsdf = divs3(bsdf, fabsf(wi.z));
  return bsdf;
}
PT_D float hair_pdf_local(const Hair &H, f3 wo_l, f3 wi_l) {
  float si = wi_l.x, ci = safe_sqrtf(1 - si * si);
  float phii = atan2f(wi_l.z, wi_l.y);
  float so = wo_l.x, co = safe_sqrtf(1 - so * so);
  float phio = atan2f(wo_l.z, wo_l.y);
  float go = asinf(clampf_(so, -1.f, 1.f));
  float etap = safe_sqrtf(sqrf(H.eta) - sqrf(so)) / co;
  float gt = asinf(clampf_(H.h / etap, -1.f, 1.f));
  float dphi = phii - phio;
  if (dphi < -PT_PI) dphi += 2.f * PT_PI;
  else if (dphi > PT_PI) dphi -= 2.f * PT_PI;
  float apdf[HAIR_PMAX + 1];
  hair_ap_pdf(H, co, apdf);
  float pdf = 0.f;
  for (int p = 0; p < HAIR_PMAX; ++p) {
    float sp, cp;
    hair_rot(H, p, so, co, sp, cp);
    pdf += apdf[p] * hair_Mp(ci, cp, si, sp, H.v[p]) * hair_Np(dphi, p, H.s, go, gt);
  }
  return pdf;
}
PT_D f3 hair_f(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) {
  Hair H = hair_params(m);
  Frame b = make_frame(h.n);
  return hair_eval(H, to_local(b, norm3(wi)), to_local(b, norm3(wo)));
}
PT_D float hair_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) {
  Hair H = hair_params(m);
  Frame b = make_frame(h.n);
  return hair_pdf_local(H, to_local(b, wo), to_local(b, norm3(wi)));
}
PT_D bool hair_scatter(const pt_material_desc &m, f3 wo, ShadePoint &h, float uc, float u0, float u1, Scatter &s) {
  Hair H = hair_params(m);
  // HairMarschner::scatter (hairMarsch.cpp:4-28): the sample and its pdf_val
  // frame on hair_tangent (the normal when unset); the attenuation is f(),
  // which frames on the normal (hairMarsch.hpp:44-54)
  const bool has_t = h.t.x != 0.f || h.t.y != 0.f || h.t.z != 0.f;
  Frame b = make_frame(has_t ? h.t : h.n);
  f3 wo_l = to_local(b, wo);
  float so = wo_l.x, co = safe_sqrtf(1 - so * so);
  float phio = atan2f(wo_l.z, wo_l.y);
  float go = asinf(clampf_(so, -1.f, 1.f));
  float apdf[HAIR_PMAX + 1];
  hair_ap_pdf(H, co, apdf);
  int p = sample_discrete<HAIR_PMAX + 1>(apdf, uc, nullptr, nullptr);
  float sp, cp;
  hair_rot(H, p, so, co, sp, cp);
  float v = H.v[p];
  float cth = 1 + v * logf(fmaxf(u0, 1e-5f) + (1.f - u0) * fast_exp(-2.f / v));
  float sth = safe_sqrtf(1 - sqrf(cth));
  float cphi = cosf(2.f * PT_PI * u1);
  float si = -cth * sp + sth * cphi * cp;
  float ci = safe_sqrtf(1 - sqrf(si));
  float etap = safe_sqrtf(sqrf(H.eta) - sqrf(so)) / co;
  float gt = asinf(clampf_(H.h / etap, -1.f, 1.f));
  float dphi = (p < HAIR_PMAX) ? hair_phi(p, go, gt) + sample_trimmed_logistic(uc, H.s, -PT_PI, PT_PI)
                               : 2.f * PT_PI * uc;
  float phii = dphi + phio;
  f3 wi = norm3(to_world(b, mk3(si, ci * cosf(phii), ci * sinf(phii))));
  if (has_t) {
    Frame bn = make_frame(h.n);
    s.att = hair_eval(H, to_local(bn, wi), to_local(bn, norm3(wo)));
  } else {
    s.att = hair_eval(H, to_local(b, wi), to_local(b, wo));
  }
  s.pdf = hair_pdf_local(H, wo_l, to_local(b, wi));
  s.wi = wi;
  s.flags = BX_GLOSSY | BX_REFL;
  return true;
}

}  // namespace pt
