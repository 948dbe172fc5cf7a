/*
 * or_math.h -- scalar vector math of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP path
 * tracer; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product never links or calls it.
 *
 * Restates gl::vec (utils/matrix.hpp:22-460) for float triples:
 *   dot      = (x*x + y*y) + z*z   (_mm_dp_ps mask 0x7F sums lanes 0..2,
 *                                    matrix.hpp:457-459)
 *   normalize divides each component by the length (matrix.hpp:265-270)
 *   cross    component order of matrix.hpp:571-579
 */
#ifndef OR_MATH_H
#define OR_MATH_H
#include <math.h>
#include <stdint.h>
#include <float.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { float x, y, z; } v3;

static inline v3 V3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline v3 vs(float s) { return V3(s, s, s); }
static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 vdivs(v3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float vlen(v3 a) { return sqrtf(vdot(a, a)); }
static inline v3 vnorm(v3 a) { float l = vlen(a); return V3(a.x / l, a.y / l, a.z / l); }
static inline v3 vcross(v3 a, v3 b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline v3 vmin3(v3 a, v3 b) { return V3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline v3 vmax3(v3 a, v3 b) { return V3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
static inline float vmaxc(v3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
static inline float sqr(float x) { return x * x; }
static inline float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
static inline float safe_sqrt(float x) { return sqrtf(fmaxf(0.f, x)); }
static inline int near_zero(v3 a) { return fabsf(a.x) < 1e-6f && fabsf(a.y) < 1e-6f && fabsf(a.z) < 1e-6f; }
static inline int v3_isfinite(v3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

/* OrthoBasis (utils/orthoBasis.hpp:15-21): w = n/|n|; a = (0,1,0) if
 * |n.x| > 0.9 else (1,0,0); v = normalize(w x a); u = w x v.              */
typedef struct { v3 u, v, w; } onb_t;
static inline onb_t onb_make(v3 n) {
  onb_t b;
  b.w = vnorm(n);
  v3 a = (fabsf(n.x) > 0.9f) ? V3(0, 1, 0) : V3(1, 0, 0);
  b.v = vnorm(vcross(b.w, a));
  b.u = vcross(b.w, b.v);
  return b;
}
/* OrthoBasis::at / toWorld (orthoBasis.hpp:23-33) */
static inline v3 onb_to_world(const onb_t *b, v3 l) {
  return vadd(vadd(vscale(b->u, l.x), vscale(b->v, l.y)), vscale(b->w, l.z));
}
/* OrthoBasis::toLocal (orthoBasis.hpp:34-36) */
static inline v3 onb_to_local(const onb_t *b, v3 w) {
  return V3(vdot(w, b->u), vdot(w, b->v), vdot(w, b->w));
}

/* gl::pbrt helpers (utils/utility.hpp:746-813) */
static inline v3 pbrt_reflect(v3 wo, v3 n) { return vadd(vneg(wo), vscale(n, 2.f * vdot(wo, n))); }
static inline float cos_theta(v3 w) { return w.z; }
static inline float cos2_theta(v3 w) { return w.z * w.z; }
static inline float abs_cos_theta(v3 w) { return fabsf(w.z); }
static inline float sin2_theta(v3 w) { return fmaxf(0.f, 1.f - cos2_theta(w)); }
static inline float sin_theta(v3 w) { return sqrtf(sin2_theta(w)); }
static inline float tan2_theta(v3 w) { return sin2_theta(w) / cos2_theta(w); }
static inline float cos_phi(v3 w) {
  float st = sin_theta(w);
  return (st == 0.f) ? 1.f : clampf(w.x / st, -1.f, 1.f);
}
static inline float sin_phi(v3 w) {
  float st = sin_theta(w);
  return (st == 0.f) ? 0.f : clampf(w.y / st, -1.f, 1.f);
}
static inline int same_hemisphere(v3 a, v3 b) { return a.z * b.z > 0.f; }
static inline v3 face_forward(v3 n, v3 v) { return (vdot(n, v) < 0.f) ? vneg(n) : n; }

/* pbrt::refract (utility.hpp:754-778) */
static inline int pbrt_refract(v3 wi, v3 n, float eta, float *etap, v3 *wt) {
  float ci = vdot(n, wi);
  if (ci < 0.f) { eta = 1.f / eta; ci = -ci; n = vneg(n); }
  float s2i = fmaxf(0.f, 1.f - sqr(ci));
  float s2t = s2i / sqr(eta);
  if (s2t >= 1.f) return 0;
  float ct = safe_sqrt(1.f - s2t);
  v3 t = vadd(vdivs(vneg(wi), eta), vscale(n, ci / eta - ct));
  *wt = vnorm(t);
  *etap = eta;
  return 1;
}

/* fresnelDielectric (material/materialMath.hpp:6-28) */
static inline float fresnel_dielectric(float ci, float eta) {
  ci = clampf(ci, -1.f, 1.f);
  if (ci < 0.f) { eta = 1.f / eta; ci = -ci; }
  float s2i = 1.f - sqr(ci);
  float s2t = s2i / sqr(eta);
  if (s2t >= 1.f) return 1.f;
  float ct = safe_sqrt(1.f - s2t);
  float rpa = (eta * ci - ct) / (eta * ci + ct);
  float rpe = (ci - eta * ct) / (ci + eta * ct);
  return (sqr(rpa) + sqr(rpe)) / 2.f;
}

#endif
