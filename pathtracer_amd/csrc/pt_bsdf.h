// pt_bsdf.h -- material evaluators on the device: Material::scatter, ::f,
// ::scatter_pdf and ::emit (material/material.hpp:15-77) for every material
// kind of the reference, in the build configuration it renders with:
// BIASED_SAMPLING (src/CMakeLists.txt:22) and TransportMode::Importance
// (config.hpp:23).
#pragma once
#include "pt_math.h"
#include "../../include/pathtracer_amd.h"

namespace pt {

enum { BX_REFL = 1, BX_TRANS = 2, BX_DIFFUSE = 4, BX_GLOSSY = 8, BX_SPECULAR = 16 };

struct ShadePoint {
  f3 n;        // normal opposing the incoming ray (HitRecord::set_normal)
  bool inside; // HitRecord::is_inside (= front face, materialBase.hpp:50)
  f3 t = {0.f, 0.f, 0.f};  // HitRecord::hair_tangent: curves set it, others leave it 0
  // a textured material's colour at the hit's texture coordinates (Lambertian
  // albedo, Disney baseColor); has_tex false: the material's own p[0..2]
  f3 tex = {0.f, 0.f, 0.f};
  bool has_tex = false;
};

struct Scatter {
  f3 wi, att;
  float pdf;
  uint32_t flags;
};

PT_D f3 mp3(const pt_material_desc &m, int o) { return mk3(m.p[o], m.p[o + 1], m.p[o + 2]); }

// cosineSampleHemiSphere (utils/random.hpp:25-31)
PT_D f3 cosine_hemi(float r1, float r2) {
  float ph = 2.f * PT_PI * r1;
  float sr = sqrtf(r2);
  return mk3(cosf(ph) * sr, sinf(ph) * sr, sqrtf(1.f - r2));
}

// ---- TrowbridgeReitzDistribution (material/materialMath.hpp:109-214) ----
struct TRD {
  float ax, ay;
};
PT_D bool trd_smooth(TRD d) { return fmaxf(d.ax, d.ay) < 1e-3f; }
PT_D float trd_D(TRD d, f3 wm) {
  float t2 = tan2t(wm);
  if (isinf(t2)) return 0.f;
  float c4 = sqrf(cos2t(wm));
  float e = t2 * (sqrf(cosphi(wm) / d.ax) + sqrf(sinphi(wm) / d.ay));
  return 1.f / (PT_PI * d.ax * d.ay * c4 * sqrf(1.f + e));
}
PT_D float trd_lambda(TRD d, f3 w) {
  float t2 = tan2t(w);
  if (isinf(t2)) return 0.f;
  float a2 = sqrf(cosphi(w) * d.ax) + sqrf(sinphi(w) * d.ay);
  return (sqrtf(1.f + a2 * t2) - 1.f) / 2.f;
}
PT_D float trd_G1(TRD d, f3 w) { return 1.f / (1.f + trd_lambda(d, w)); }
PT_D float trd_G(TRD d, f3 wo, f3 wi) { return 1.f / (1.f + trd_lambda(d, wo) + trd_lambda(d, wi)); }
PT_D float trd_Dw(TRD d, f3 w, f3 wm) { return trd_G1(d, w) / abs_cost(w) * trd_D(d, wm) * fabsf(dot3(w, wm)); }
// Visible-normal sampling (Heitz 2018) as in sample_wm (materialMath.hpp:178-200)
PT_D f3 trd_sample_wm(TRD d, f3 w, float u0, float u1) {
  f3 wh = norm3(mk3(d.ax * w.x, d.ay * w.y, w.z));
  if (wh.z < 0.f) wh = neg3(wh);
  f3 T1 = (wh.z < 0.99999f) ? norm3(cross3(mk3(0.f, 0.f, 1.f), wh)) : mk3(1.f, 0.f, 0.f);
  f3 T2 = cross3(wh, T1);
  float r = sqrtf(u0);
  float th = 2.f * PT_PI * u1;
  float px = r * cosf(th), py = r * sinf(th);
  float h = sqrtf(1.f - sqrf(px));
  float t = (1.f + wh.z) / 2.f;
  py = h + t * (py - h);
  float pl = sqrtf(px * px + py * py);
  float pz = sqrtf(fmaxf(0.f, 1.f - sqrf(pl)));
  f3 nh = add3(add3(scl3(T1, px), scl3(T2, py)), scl3(wh, pz));
  return norm3(mk3(d.ax * nh.x, d.ay * nh.y, fmaxf(1e-6f, nh.z)));
}

// fresnelComplex (materialMath.hpp:30-53) written out in real arithmetic:
// eta = a + ib; z = 1 - sin2/eta^2; cos_t = sqrt(z) (principal branch).
PT_D float fresnel_complex1(float ci, float a, float b) {
  ci = clampf_(ci, 0.f, 1.f);
  float s2i = 1.f - sqrf(ci);
  // eta^2 = (a^2 - b^2) + i 2ab ; s2i / eta^2
  float er = a * a - b * b, ei = 2.f * a * b;
  float den = er * er + ei * ei;
  float qr = s2i * er / den, qi = -s2i * ei / den;
  float zr = 1.f - qr, zi = -qi;
  // principal sqrt of (zr + i zi)
  float mag = sqrtf(zr * zr + zi * zi);
  float cr = sqrtf(fmaxf(0.f, 0.5f * (mag + zr)));
  float cim = copysignf(sqrtf(fmaxf(0.f, 0.5f * (mag - zr))), zi);
  // r_parl = (eta ci - ct)/(eta ci + ct); r_perp = (ci - eta ct)/(ci + eta ct)
  float nr = a * ci - cr, ni = b * ci - cim;
  float dr = a * ci + cr, di = b * ci + cim;
  float rpa = (nr * nr + ni * ni) / (dr * dr + di * di);
  float ecr = a * cr - b * cim, eci = a * cim + b * cr;  // eta * ct
  float n2r = ci - ecr, n2i = -eci, d2r = ci + ecr, d2i = eci;
  float rpe = (n2r * n2r + n2i * n2i) / (d2r * d2r + d2i * d2i);
  return (rpa + rpe) / 2.f;
}
PT_D f3 fresnel_complex(float ci, f3 eta, f3 k) {
  return mk3(fresnel_complex1(ci, eta.x, k.x), fresnel_complex1(ci, eta.y, k.y),
             fresnel_complex1(ci, eta.z, k.z));
}

// ---- Lambertian (material.hpp:79-137) ----
PT_D float cosine_pdf(f3 n, f3 wi) {
  Frame b = make_frame(n);
  return fmaxf(dot3(norm3(wi), b.w), 0.f) / PT_PI;
}
// CosinePDF(n)::get (pdf.hpp:19-37), for integrators that fall back to it
PT_D f3 cosine_dir(f3 n, float u0, float u1) {
  Frame b = make_frame(n);
  return norm3(to_world(b, cosine_hemi(u0, u1)));
}
PT_D bool lambert_scatter(const pt_material_desc &m, const ShadePoint &h, float u0, float u1, Scatter &s) {
  Frame b = make_frame(h.n);
  s.flags = BX_DIFFUSE | BX_REFL;
  s.att = scl3(h.has_tex ? h.tex : mp3(m, 0), 1.f / PT_PI);
  s.wi = norm3(to_world(b, cosine_hemi(u0, u1)));
  s.pdf = fmaxf(dot3(norm3(s.wi), b.w), 0.f) / PT_PI;
  return true;
}

// ---- Conductor (conductor.hpp:4-145, MicrofacetPDF mfPDF.hpp:6-62) ----
PT_D float mf_pdf_at(TRD d, const Frame &b, f3 wo_l, f3 wi_world) {
  f3 wi = to_local(b, norm3(wi_world));
  if (wi.z <= 0.f) return 0.f;
  f3 m = norm3(add3(wo_l, wi));
  float den = 4.f * fabsf(dot3(wo_l, m));
  if (den < 1e-6f) return 0.f;
  return trd_Dw(d, wo_l, m) / den;
}
PT_D f3 conductor_f(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) {
  TRD d{m.p[6], m.p[7]};
  Frame b = make_frame(h.n);
  f3 wo_l = to_local(b, norm3(wo)), wi_l = to_local(b, norm3(wi));
  if (wi_l.z <= 0.f || wo_l.z <= 0.f) return s3(0.f);
  if (trd_smooth(d)) return s3(0.f);
  f3 ml = norm3(add3(wo_l, wi_l));
  float D = trd_D(d, ml), G = trd_G(d, wo_l, wi_l);
  f3 F = fresnel_complex(fabsf(dot3(wo_l, ml)), mp3(m, 0), mp3(m, 3));
  float den = 4.f * abs_cost(wo_l) * abs_cost(wi_l);
  return divs3(scl3(scl3(F, D), G), den);
}
PT_D bool conductor_scatter(const pt_material_desc &m, f3 wo, const ShadePoint &h, float u0, float u1,
                            Scatter &s) {
  TRD d{m.p[6], m.p[7]};
  if (trd_smooth(d)) {
    f3 wi = norm3(reflect_pbrt(wo, h.n));
    if (dot3(h.n, wi) <= 0.f) return false;
    s.wi = wi;
    s.att = fresnel_complex(fabsf(dot3(h.n, wi)), mp3(m, 0), mp3(m, 3));
    s.flags = BX_SPECULAR | BX_REFL;
    s.pdf = 0.f;
    return true;
  }
  Frame b = make_frame(h.n);
  f3 wo_l = to_local(b, norm3(wo));
  f3 ml = trd_sample_wm(d, wo_l, u0, u1);
  f3 wi_l = reflect_pbrt(wo_l, ml);
  if (!same_hemi(wo_l, wi_l)) wi_l.z = -wo_l.z;  // BIASED_SAMPLING (mfPDF.hpp:28-30)
  f3 wi = norm3(to_world(b, wi_l));
  if (near_zero3(wi)) return false;
  if (dot3(h.n, wi) <= 0.f) return false;
  s.wi = wi;
  s.flags = BX_GLOSSY | BX_REFL;
  s.att = conductor_f(m, wo, wi, h);
  s.pdf = mf_pdf_at(d, b, wo_l, wi);
  return true;
}
PT_D float conductor_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) {
  TRD d{m.p[6], m.p[7]};
  if (trd_smooth(d)) return 0.f;
  Frame b = make_frame(h.n);
  f3 wo_l = to_local(b, norm3(wo));
  if (dot3(h.n, wi) <= 0.f) return 0.f;
  return mf_pdf_at(d, b, wo_l, wi);
}

// conductor_f and conductor_pdf in one pass: the frame, the local directions,
// the half vector, D and Lambda(wo) are computed once; each output keeps its
// own early-outs and the same operations in the same order, so both equal
// the separate functions' bit for bit.
PT_D void conductor_f_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h, f3 &f, float &pdf) {
  f = s3(0.f);
  pdf = 0.f;
  TRD d{m.p[6], m.p[7]};
  if (trd_smooth(d)) return;
  Frame b = make_frame(h.n);
  f3 wo_l = to_local(b, norm3(wo)), wi_l = to_local(b, norm3(wi));
  if (wi_l.z <= 0.f) return;  // both outputs are 0
  f3 ml = norm3(add3(wo_l, wi_l));
  const float D = trd_D(d, ml), lo = trd_lambda(d, wo_l);
  if (wo_l.z > 0.f) {
    const float G = 1.f / (1.f + lo + trd_lambda(d, wi_l));
    f3 F = fresnel_complex(fabsf(dot3(wo_l, ml)), mp3(m, 0), mp3(m, 3));
    float den = 4.f * abs_cost(wo_l) * abs_cost(wi_l);
    f = divs3(scl3(scl3(F, D), G), den);
  }
  if (dot3(h.n, wi) <= 0.f) return;
  const float den = 4.f * fabsf(dot3(wo_l, ml));
  if (den < 1e-6f) return;
  pdf = 1.f / (1.f + lo) / abs_cost(wo_l) * D * fabsf(dot3(wo_l, ml)) / den;
}

// ---- MFDielectric (dielectric.hpp:5-167, MFDielectricPDF mfDielectricPDF.hpp:7-147) ----
PT_D f3 mfdiel_f(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) {
  float eta = m.p[0];
  TRD d{m.p[1], m.p[2]};
  if (eta == 1.f || trd_smooth(d)) return s3(0.f);
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
  if (refl) return s3(trd_D(d, wm) * trd_G(d, wo_l, wi_l) * F / fabsf(4.f * ci * co));
  float den = sqrf(dot3(wi_l, wm) + dot3(wo_l, wm) / etap) * ci * co;
  float ft = trd_D(d, wm) * (1.f - F) * trd_G(d, wo_l, wi_l) * fabsf(dot3(wi_l, wm) * dot3(wo_l, wm) / den);
  return s3(ft);
}
PT_D float mfdiel_pdf_local(float eta, TRD d, f3 wo, f3 wi) {
  if (eta == 1.f || trd_smooth(d)) return 0.f;
  if (wi.z == 0.f) return 0.f;
  float co = wo.z, ci = wi.z;
  bool refl = ci * co > 0.f;
  float etap = 1.f;
  if (!refl) etap = co > 0.f ? eta : (1.f / eta);
  f3 wm = add3(scl3(wi, etap), wo);
  if (ci == 0.f || co == 0.f || len3(wm) == 0.f) return 0.f;
  wm = face_fwd(norm3(wm), mk3(0.f, 0.f, 1.f));
  if (dot3(wm, wi) * ci < 0.f || dot3(wm, wo) * wo.z < 0.f) return 0.f;
  float R = fresnel_diel(dot3(wo, wm), eta), T = 1.f - R;
  if (R + T == 0.f) return 0.f;
  if (refl) return trd_Dw(d, wo, wm) / (4.f * fabsf(dot3(wo, wm))) * R / (R + T);
  float den = sqrf(dot3(wi, wm) + dot3(wo, wm) / etap);
  float dwm = fabsf(dot3(wi, wm)) / den;
  return trd_Dw(d, wo, wm) * dwm * T / (R + T);
}
// mfdiel_f and mfdiel_pdf in one pass (shared frame, local directions,
// half vector, Fresnel term, D and Lambda(wo)); both outputs equal the
// separate functions' bit for bit.
PT_D void mfdiel_f_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h, f3 &f, float &pdf) {
  f = s3(0.f);
  pdf = 0.f;
  float eta = m.p[0];
  TRD d{m.p[1], m.p[2]};
  if (eta == 1.f || trd_smooth(d)) return;
  Frame b = make_frame(h.n);
  f3 wo_l = to_local(b, norm3(wo)), wi_l = to_local(b, norm3(wi));
  float co = wo_l.z, ci = wi_l.z;
  bool refl = ci * co > 0.f;
  float etap = 1.f;
  if (!refl) etap = co > 0.f ? eta : (1.f / eta);
  f3 wm = add3(scl3(wi_l, etap), wo_l);
  if (ci == 0.f || co == 0.f || len3(wm) == 0.f) return;
  wm = face_fwd(norm3(wm), mk3(0.f, 0.f, 1.f));
  if (dot3(wm, wi_l) * ci < 0.f || dot3(wm, wo_l) * co < 0.f) return;
  const float F = fresnel_diel(dot3(wo_l, wm), eta), T = 1.f - F;
  const float D = trd_D(d, wm), lo = trd_lambda(d, wo_l);
  const float G = 1.f / (1.f + lo + trd_lambda(d, wi_l));
  const float Dw = 1.f / (1.f + lo) / abs_cost(wo_l) * D * fabsf(dot3(wo_l, wm));
  if (refl) {
    f = s3(D * G * F / fabsf(4.f * ci * co));
    if (F + T != 0.f) pdf = Dw / (4.f * fabsf(dot3(wo_l, wm))) * F / (F + T);
    return;
  }
  const float den = sqrf(dot3(wi_l, wm) + dot3(wo_l, wm) / etap) * ci * co;
  f = s3(D * (1.f - F) * G * fabsf(dot3(wi_l, wm) * dot3(wo_l, wm) / den));
  if (F + T != 0.f) {
    const float den_p = sqrf(dot3(wi_l, wm) + dot3(wo_l, wm) / etap);
    pdf = Dw * (fabsf(dot3(wi_l, wm)) / den_p) * T / (F + T);
  }
}
PT_D bool mfdiel_scatter(const pt_material_desc &m, f3 wo, const ShadePoint &h, float uc, float u0,
                         float u1, Scatter &s) {
  float eta = m.p[0];
  TRD d{m.p[1], m.p[2]};
  if (eta == 1.f || trd_smooth(d)) {
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
    s.pdf = T;
    s.att = s3(1.f);
    return true;
  }
  Frame b = make_frame(h.n);
  f3 wo_l = to_local(b, norm3(wo));
  f3 ml = trd_sample_wm(d, wo_l, u0, u1);
  float R = fresnel_diel(dot3(wo_l, ml), eta), T = 1.f - R;
  if (R + T == 0.f) return false;
  f3 wi_l;
  if (uc < R / (R + T)) {
    wi_l = reflect_pbrt(wo_l, ml);
    if (!same_hemi(wo_l, wi_l)) wi_l.z = -wi_l.z;  // BIASED_SAMPLING (mfDielectricPDF.hpp:47-49)
  } else {
    float es;
    bool valid = refract_pbrt(wo_l, ml, eta, &es, &wi_l);
    if (!valid || same_hemi(wo_l, wi_l)) wi_l = reflect_pbrt(wo_l, ml);  // (:70-72)
  }
  f3 wi = to_world(b, wi_l);
  if (near_zero3(wi)) return false;
  bool is_refl = dot3(to_local(b, wi), to_local(b, wo)) > 0.f;
  s.wi = wi;
  s.flags = BX_GLOSSY | (is_refl ? BX_REFL : BX_TRANS);
  mfdiel_f_pdf(m, wo, wi, h, s.att, s.pdf);
  return true;
}
PT_D float mfdiel_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) {
  float eta = m.p[0];
  TRD d{m.p[1], m.p[2]};
  if (eta == 1.f || trd_smooth(d)) return 0.f;
  Frame b = make_frame(h.n);
  return mfdiel_pdf_local(eta, d, to_local(b, norm3(wo)), to_local(b, norm3(wi)));
}

// ---- ThinDielectric (thinDielectric.hpp:3-70); the split-ray variant is
// taken as a choice with probability R, which has the same expectation as
// tracing both branches (maxdepth_nee.hpp:51-71). ----
PT_D bool thin_scatter(const pt_material_desc &m, f3 wo, const ShadePoint &h, float uc, Scatter &s) {
  float eta = m.p[0];
  float aco = fabsf(dot3(h.n, wo));
  float R = fresnel_diel(aco, eta), T = 1.f - R;
  if (R < 1.f) {
    R += sqrf(T) * R / (1.f - sqrf(R));
    T = 1.f - R;
  }
  if (R == 0.f && T == 0.f) return false;
  if (uc < R / (R + T)) {
    s.wi = norm3(reflect_pbrt(wo, h.n));
    s.flags = BX_SPECULAR | BX_REFL;
    s.pdf = R;
  } else {
    s.wi = neg3(norm3(wo));
    s.flags = BX_SPECULAR | BX_TRANS;
    s.pdf = T;
  }
  s.att = s3(1.f);
  return true;
}

// ---- Phong (material.hpp:303-465, PhongLobePDF pdf.hpp:57-91) ----
PT_D f3 gl_reflect(f3 v, f3 n) {  // gl::reflect (utility.hpp:39-45)
  f3 vv = norm3(v), nn = norm3(n);
  float dt = dot3(vv, nn);
  return norm3(sub3(vv, scl3(nn, 2.f * dt)));
}
PT_D float phong_lobe_pdf(f3 R, float shin, f3 dir) {
  float c = fmaxf(dot3(norm3(dir), norm3(R)), 0.f);
  return (shin + 1.f) / (2.f * PT_PI) * powf(c, shin);
}
PT_D bool phong_scatter(const pt_material_desc &m, f3 wo, const ShadePoint &h, float u0, float u1,
                        float aux, Scatter &s) {
  f3 dif = mp3(m, 0), spe = mp3(m, 3);
  float shin = m.p[9];
  float si = (spe.x + spe.y + spe.z) / 3.f, di = (dif.x + dif.y + dif.z) / 3.f;
  float sp = si / (si + di);
  if (aux < sp) {
    f3 R = gl_reflect(neg3(wo), h.n);
    Frame b = make_frame(R);
    float th = acosf(powf(u0, 1.f / (shin + 1.f)));
    float ph = 2.f * PT_PI * u1;
    float st = sinf(th);
    f3 dir = norm3(to_world(b, mk3(st * cosf(ph), st * sinf(ph), cosf(th))));
    s.flags = BX_GLOSSY | BX_REFL;
    s.wi = dir;
    s.att = spe;
    s.pdf = phong_lobe_pdf(norm3(R), shin, dir);
    return true;
  }
  Frame b = make_frame(h.n);
  s.flags = BX_DIFFUSE | BX_REFL;
  s.att = dif;
  s.wi = to_world(b, cosine_hemi(u0, u1));
  s.pdf = 0.f;  // never set by the reference's diffuse branch (material.hpp:357-365)
  return true;
}
PT_D f3 phong_f(const pt_material_desc &m, f3 wo_w, f3 wi_w, const ShadePoint &h) {
  f3 dif = mp3(m, 0), spe = mp3(m, 3);
  float shin = m.p[9];
  f3 wo = norm3(wo_w), wi = norm3(wi_w);
  float cti = dot3(h.n, wi);
  if (cti <= 0.f) return s3(0.f);
  float si = (spe.x + spe.y + spe.z) / 3.f, di = (dif.x + dif.y + dif.z) / 3.f;
  float sp = si / (si + di);
  f3 db = scl3(dif, 1.f / PT_PI);
  f3 R = reflect_pbrt(wo, h.n);
  float pt = powf(fmaxf(dot3(wi, R), 0.f), shin);
  f3 sb = scl3(spe, (shin + 2.f) / (2.f * PT_PI) * pt);
  return add3(scl3(sb, sp), scl3(db, 1.f - sp));
}
PT_D float phong_pdf(const pt_material_desc &m, f3 wo_w, f3 wi_w, const ShadePoint &h) {
  f3 dif = mp3(m, 0), spe = mp3(m, 3);
  float shin = m.p[9];
  f3 wo = norm3(wo_w), wi = norm3(wi_w);
  float si = (spe.x + spe.y + spe.z) / 3.f, di = (dif.x + dif.y + dif.z) / 3.f;
  float tot = si + di;
  if (tot == 0.f) return 0.f;
  float ps = si / tot, pd = di / tot, pl = 0.f, pdl = 0.f;
  if (ps > 0.f) pl = phong_lobe_pdf(gl_reflect(neg3(wo), norm3(h.n)), shin, wi);
  if (pd > 0.f) {
    float c = dot3(wi, norm3(h.n));
    if (c > 0.f) pdl = c / PT_PI;
  }
  return ps * pl + pd * pdl;
}

// ---- Kajiya-Kay (no reference code; README.md:74 lists it, TODO.md:20 is
// open).  Kajiya & Kay (1989) about the fibre tangent t: the hair tangent of
// a curve hit, else the first axis of OrthoBasis(n) (a direction in the
// surface, as for a hair card):
//   f = kd/pi sin(t,wi) + ks (p+2)/(2pi) max(0, sin(t,wi) sin(t,wo) - cos(t,wi) cos(t,wo))^p,
// whose specular term peaks on the mirror cone t.wo = -t.wi.  Sampled by the
// cosine lobe about n. ----
PT_D f3 kk_tangent(const ShadePoint &h) {
  if (!near_zero3(h.t)) return norm3(h.t);
  return make_frame(h.n).u;
}
PT_D f3 kk_f(const pt_material_desc &m, f3 wo_w, f3 wi_w, const ShadePoint &h) {
  f3 t = kk_tangent(h), wo = norm3(wo_w), wi = norm3(wi_w);
  float ci = dot3(t, wi), co = dot3(t, wo);
  float si = sqrtf(fmaxf(0.f, 1.f - ci * ci)), so = sqrtf(fmaxf(0.f, 1.f - co * co));
  float ex = m.p[6];
  float spec = powf(fmaxf(0.f, si * so - ci * co), ex) * ((ex + 2.f) / (2.f * PT_PI));
  return add3(scl3(mp3(m, 0), si / PT_PI), scl3(mp3(m, 3), spec));
}
PT_D bool kk_scatter(const pt_material_desc &m, f3 wo, const ShadePoint &h, float u0, float u1, Scatter &s) {
  Frame b = make_frame(h.n);
  s.wi = norm3(to_world(b, cosine_hemi(u0, u1)));
  s.pdf = cosine_pdf(h.n, s.wi);
  if (s.pdf <= 0.f) return false;
  s.flags = BX_GLOSSY | BX_REFL;
  s.att = kk_f(m, wo, s.wi, h);
  return true;
}

}  // namespace pt

#include "pt_bsdf_disney.h"

namespace pt {

// ---- dispatch ----
// MS is the set of material kinds compiled into the caller (bit k = kind k):
// a kernel instantiated for a scene's kinds carries no register pressure from
// the materials that scene does not use.  Kinds outside MS scatter nothing.
constexpr uint32_t kMatAll = 0xffffffffu;
// BSDFs that only reflect: their f (Lambertian's in particular,
// material.hpp:128-133) is not zero behind the surface, so estimators that
// use |cos| must exclude those directions.
PT_HD bool bsdf_reflect_only(uint32_t kind) {
  return kind == PT_MAT_LAMBERTIAN || kind == PT_MAT_CONDUCTOR || kind == PT_MAT_PHONG;
}

constexpr uint32_t kMatBasic = (1u << PT_MAT_LAMBERTIAN) | (1u << PT_MAT_EMITTER) | (1u << PT_MAT_CONDUCTOR) |
                               (1u << PT_MAT_MFDIELECTRIC) | (1u << PT_MAT_THINDIELECTRIC) |
                               (1u << PT_MAT_PHONG) | (1u << PT_MAT_KAJIYA_KAY);
// The smallest set: Lambertian walls, emitters and the microfacet conductor
// and dielectric (the Cornell and glass/HDRI scenes).
constexpr uint32_t kMatMicro = (1u << PT_MAT_LAMBERTIAN) | (1u << PT_MAT_EMITTER) | (1u << PT_MAT_CONDUCTOR) |
                               (1u << PT_MAT_MFDIELECTRIC);
// Material-set bit of scenes with a textured material (only the all-kinds
// set carries the texture lookups).
constexpr uint32_t kMatTextured = 1u << 31;
#define PT_IN(MS, K) (((MS) >> (K)) & 1u)

template <uint32_t MS = kMatAll>
PT_D bool bsdf_scatter(const pt_material_desc &m, f3 wo, ShadePoint &h, float uc, float u0, float u1,
                       float aux, Scatter &s) {
  s.flags = 0;
  s.pdf = 0.f;
  s.att = s3(0.f);
  s.wi = s3(0.f);
  const uint32_t k = m.kind;
  if (PT_IN(MS, PT_MAT_LAMBERTIAN) && k == PT_MAT_LAMBERTIAN) return lambert_scatter(m, h, u0, u1, s);
  if (PT_IN(MS, PT_MAT_CONDUCTOR) && k == PT_MAT_CONDUCTOR) return conductor_scatter(m, wo, h, u0, u1, s);
  if (PT_IN(MS, PT_MAT_MFDIELECTRIC) && k == PT_MAT_MFDIELECTRIC) return mfdiel_scatter(m, wo, h, uc, u0, u1, s);
  if (PT_IN(MS, PT_MAT_THINDIELECTRIC) && k == PT_MAT_THINDIELECTRIC) return thin_scatter(m, wo, h, uc, s);
  if (PT_IN(MS, PT_MAT_PHONG) && k == PT_MAT_PHONG) return phong_scatter(m, wo, h, u0, u1, aux, s);
  if (PT_IN(MS, PT_MAT_DISNEY) && k == PT_MAT_DISNEY) return disney_scatter(m, wo, h, uc, u0, u1, s);
  if (PT_IN(MS, PT_MAT_HAIR) && k == PT_MAT_HAIR) return hair_scatter(m, wo, h, uc, u0, u1, s);
  if (PT_IN(MS, PT_MAT_KAJIYA_KAY) && k == PT_MAT_KAJIYA_KAY) return kk_scatter(m, wo, h, u0, u1, s);
  return false;
}
template <uint32_t MS = kMatAll>
PT_D f3 bsdf_f(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) {
  const uint32_t k = m.kind;
  if (PT_IN(MS, PT_MAT_LAMBERTIAN) && k == PT_MAT_LAMBERTIAN) return scl3(h.has_tex ? h.tex : mp3(m, 0), 1.f / PT_PI);
  if (PT_IN(MS, PT_MAT_CONDUCTOR) && k == PT_MAT_CONDUCTOR) return conductor_f(m, wo, wi, h);
  if (PT_IN(MS, PT_MAT_MFDIELECTRIC) && k == PT_MAT_MFDIELECTRIC) return mfdiel_f(m, wo, wi, h);
  if (PT_IN(MS, PT_MAT_PHONG) && k == PT_MAT_PHONG) return phong_f(m, wo, wi, h);
  if (PT_IN(MS, PT_MAT_DISNEY) && k == PT_MAT_DISNEY) return disney_f(m, wo, wi, h);
  if (PT_IN(MS, PT_MAT_HAIR) && k == PT_MAT_HAIR) return hair_f(m, wo, wi, h);
  if (PT_IN(MS, PT_MAT_KAJIYA_KAY) && k == PT_MAT_KAJIYA_KAY) return kk_f(m, wo, wi, h);
  return s3(0.f);
}
template <uint32_t MS = kMatAll>
PT_D float bsdf_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) {
  const uint32_t k = m.kind;
  if (PT_IN(MS, PT_MAT_LAMBERTIAN) && k == PT_MAT_LAMBERTIAN) return cosine_pdf(h.n, wi);
  if (PT_IN(MS, PT_MAT_CONDUCTOR) && k == PT_MAT_CONDUCTOR) return conductor_pdf(m, wo, wi, h);
  if (PT_IN(MS, PT_MAT_MFDIELECTRIC) && k == PT_MAT_MFDIELECTRIC) return mfdiel_pdf(m, wo, wi, h);
  if (PT_IN(MS, PT_MAT_PHONG) && k == PT_MAT_PHONG) return phong_pdf(m, wo, wi, h);
  if (PT_IN(MS, PT_MAT_DISNEY) && k == PT_MAT_DISNEY) return disney_pdf(m, wo, wi, h);
  if (PT_IN(MS, PT_MAT_HAIR) && k == PT_MAT_HAIR) return hair_pdf(m, wo, wi, h);
  if (PT_IN(MS, PT_MAT_KAJIYA_KAY) && k == PT_MAT_KAJIYA_KAY) return cosine_pdf(h.n, wi);
  return 0.f;
}
// bsdf_f and bsdf_pdf of one direction, fused where the kinds share terms
// (the results equal the two calls').
template <uint32_t MS = kMatAll>
PT_D void bsdf_f_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h, f3 &f, float &pdf) {
  const uint32_t k = m.kind;
  if (PT_IN(MS, PT_MAT_CONDUCTOR) && k == PT_MAT_CONDUCTOR) {
    conductor_f_pdf(m, wo, wi, h, f, pdf);
  } else if (PT_IN(MS, PT_MAT_MFDIELECTRIC) && k == PT_MAT_MFDIELECTRIC) {
    mfdiel_f_pdf(m, wo, wi, h, f, pdf);
  } else if (PT_IN(MS, PT_MAT_DISNEY) && k == PT_MAT_DISNEY) {
    disney_f_pdf(m, wo, wi, h, f, pdf);
  } else {
    f = bsdf_f<MS>(m, wo, wi, h);
    pdf = bsdf_pdf<MS>(m, wo, wi, h);
  }
}
// DiffuseEmitter::emit (material.hpp:487-493)
PT_D f3 bsdf_emit(const pt_material_desc &m, f3 ray_d, const ShadePoint &h) {
  if (m.kind != PT_MAT_EMITTER) return s3(0.f);
  if (dot3(h.n, neg3(ray_d)) > 0.f) return scl3(mp3(m, 0), m.p[3]);
  return s3(0.f);
}

}  // namespace pt
