// pt_math.h -- device-side vector math for the gfx950 path tracer.
//
// Float triples are kept as three scalars (no float4 padding) so that SoA
// buffers and registers hold exactly what the math needs.  Operation order
// follows gl::vec (utils/matrix.hpp:22-460): dot = (x*x + y*y) + z*z,
// normalize divides each component by the length.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PT_HD __host__ __device__ __forceinline__
#define PT_D __device__ __forceinline__

#ifndef PT_PI
#define PT_PI 3.14159265358979323846f
#endif

struct f3 {
  float x, y, z;
};

// Division and square root whose results must be bit-identical to IEEE
// arithmetic (sampler values, camera ray set-up, primitive hit decisions).
// The kernels are compiled with correctly rounded division and square root,
// so these are the plain operators; they mark the places that rely on it.
PT_HD float div_rn(float a, float b) { return a / b; }
PT_HD float rcp_rn(float a) { return 1.0f / a; }
PT_HD float sqrt_rn(float a) { return sqrtf(a); }
// Hardware reciprocal (about 1 ulp): for the BVH slab test only.
PT_HD float rcp_fast(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(a);
#else
  return 1.0f / a;
#endif
}

PT_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
PT_HD f3 s3(float s) { return mk3(s, s, s); }
PT_HD f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
PT_HD f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
PT_HD f3 mul3(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
PT_HD f3 scl3(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
PT_HD f3 divs3(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
PT_HD f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }
PT_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PT_HD float len3(f3 a) { return sqrtf(dot3(a, a)); }
PT_HD f3 norm3(f3 a) { float l = len3(a); return mk3(a.x / l, a.y / l, a.z / l); }
PT_HD f3 norm3_rn(f3 a) {
  float l = sqrt_rn(dot3(a, a));
  return mk3(div_rn(a.x, l), div_rn(a.y, l), div_rn(a.z, l));
}
PT_HD f3 cross3(f3 a, f3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
PT_HD float comp3(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
PT_HD float maxc3(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
PT_HD float sqrf(float x) { return x * x; }
PT_HD float clampf_(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
PT_HD float safe_sqrtf(float x) { return sqrtf(fmaxf(0.f, x)); }
PT_HD bool near_zero3(f3 a) { return fabsf(a.x) < 1e-6f && fabsf(a.y) < 1e-6f && fabsf(a.z) < 1e-6f; }
PT_HD bool finite3(f3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

// Orthonormal frame of OrthoBasis (utils/orthoBasis.hpp:15-21): w = n/|n|,
// helper axis (0,1,0) when |n.x| > 0.9 else (1,0,0), v = normalize(w x a),
// u = w x v.
struct Frame {
  f3 u, v, w;
};
PT_HD Frame make_frame(f3 n) {
  Frame b;
  b.w = norm3(n);
  f3 a = (fabsf(n.x) > 0.9f) ? mk3(0.f, 1.f, 0.f) : mk3(1.f, 0.f, 0.f);
  b.v = norm3(cross3(b.w, a));
  b.u = cross3(b.w, b.v);
  return b;
}
PT_HD f3 to_world(const Frame &b, f3 l) {
  return add3(add3(scl3(b.u, l.x), scl3(b.v, l.y)), scl3(b.w, l.z));
}
PT_HD f3 to_local(const Frame &b, f3 w) { return mk3(dot3(w, b.u), dot3(w, b.v), dot3(w, b.w)); }

// Local-frame trigonometry (gl::pbrt, utils/utility.hpp:780-813)
PT_HD f3 reflect_pbrt(f3 wo, f3 n) { return add3(neg3(wo), scl3(n, 2.f * dot3(wo, n))); }
PT_HD float cos2t(f3 w) { return w.z * w.z; }
PT_HD float abs_cost(f3 w) { return fabsf(w.z); }
PT_HD float sin2t(f3 w) { return fmaxf(0.f, 1.f - cos2t(w)); }
PT_HD float sint(f3 w) { return sqrtf(sin2t(w)); }
PT_HD float tan2t(f3 w) { return sin2t(w) / cos2t(w); }
PT_HD float cosphi(f3 w) {
  float st = sint(w);
  return (st == 0.f) ? 1.f : clampf_(w.x / st, -1.f, 1.f);
}
PT_HD float sinphi(f3 w) {
  float st = sint(w);
  return (st == 0.f) ? 0.f : clampf_(w.y / st, -1.f, 1.f);
}
PT_HD bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0.f; }
PT_HD f3 face_fwd(f3 n, f3 v) { return (dot3(n, v) < 0.f) ? neg3(n) : n; }

// pbrt::refract (utility.hpp:754-778)
PT_HD bool refract_pbrt(f3 wi, f3 n, float eta, float *etap, f3 *wt) {
  float ci = dot3(n, wi);
  if (ci < 0.f) { eta = 1.f / eta; ci = -ci; n = neg3(n); }
  float s2i = fmaxf(0.f, 1.f - sqrf(ci));
  float s2t = s2i / sqrf(eta);
  if (s2t >= 1.f) return false;
  float ct = safe_sqrtf(1.f - s2t);
  f3 t = add3(divs3(neg3(wi), eta), scl3(n, ci / eta - ct));
  *wt = norm3(t);
  *etap = eta;
  return true;
}

// fresnelDielectric (material/materialMath.hpp:6-28)
PT_HD float fresnel_diel(float ci, float eta) {
  ci = clampf_(ci, -1.f, 1.f);
  if (ci < 0.f) { eta = 1.f / eta; ci = -ci; }
  float s2i = 1.f - sqrf(ci);
  float s2t = s2i / sqrf(eta);
  if (s2t >= 1.f) return 1.f;
  float ct = safe_sqrtf(1.f - s2t);
  float rpa = (eta * ci - ct) / (eta * ci + ct);
  float rpe = (ci - eta * ct) / (ci + eta * ct);
  return (sqrf(rpa) + sqrf(rpe)) / 2.f;
}

// ---- counter-based random stream (the reference's gl::rand_num is an
// mt19937 seeded from std::random_device; a keyed hash makes every path
// sample reproducible).  lowbias32 integer hash (C. Wellons). ----
PT_HD uint32_t pt_hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
PT_HD uint32_t pt_rng_u32(uint32_t a, uint32_t b, uint32_t c) {
  return pt_hash32(a ^ pt_hash32(b ^ pt_hash32(c)));
}
PT_HD float pt_rng(uint32_t a, uint32_t b, uint32_t c) {
  return (float)(pt_rng_u32(a, b, c) >> 8) * 0x1p-24f;
}
PT_HD uint32_t pt_pixel_seed(uint32_t pixel, uint32_t seed) {
  return pt_hash32(pixel ^ pt_hash32(seed + 0x9e3779b9u));
}

// Random-stream slots: slot = vertex*64 + purpose.
enum {
  SLOT_LENS_X = 0, SLOT_LENS_Y = 1, SLOT_TIME = 2, SLOT_AUX = 4, SLOT_RR = 5,
  SLOT_MIS_LIGHT = 6, SLOT_MIX = 7, SLOT_NEE_OFF = 8, SLOT_RIS = 16, SLOT_NEE_LIGHT = 40, SLOT_GUIDE = 44
};
PT_HD uint32_t rslot(uint32_t d, uint32_t purpose) { return d * 64u + purpose; }
