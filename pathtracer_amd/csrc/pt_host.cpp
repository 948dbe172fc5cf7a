// pt_host.cpp -- host side of the MI355X path tracer: scene assembly, BVH
// build, light discovery, device upload, wavefront orchestration and the C
// ABI of include/pathtracer_amd.h.
//
// Reference interfaces replaced (file:line):
//   SceneInfo::render                 render/renderManager.hpp:51-158
//   ObjectList / BVHNode              base/objectList.hpp:9-150, primitives/bvh.hpp:8-102
//   TriangleMesh / MeshBVHNode        primitives/triangleMesh.cpp:5-56, meshBVH.cpp:6-143
//   discover_emissive_objects_as_lights  base/lightDiscovery.hpp:5-81
//   FrameBuffer offsets               base/framebuffer.cpp:131-145
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pathtracer_amd.h"
#include "pt_guide_host.h"
#include "pt_bsdf.h"
#include "pt_launch.h"
#include "pt_lbvh.h"

using namespace pt;

#include "pt_common.h"

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;
pt_status pt_fail(pt_status st, const std::string &msg) {
  g_err = msg;
  return st;
}

// Wavefront state of one render stream (a scene's own renders, or a batch).
struct WaveStore {
  DevBuf buf, counts, stats, res;
  Wave wave{};
  uint32_t cap = 0;
};

// ------------------------------------------------------------------ host-side scene records
struct HTri { f3 p[3]; uint32_t mesh; int smooth; };  // smooth: index into pt_scene::normals3 or -1
struct HMesh {
  uint32_t first, count, mat, flags, seq;
  uint32_t nverts = 0;
  std::vector<int32_t> idx;  // the index buffer, for position updates
  f3 box_lo, box_hi;          // TriangleMesh::meshAABB: vertex bounds -/+ gl::epsilon (triangleMesh.cpp:10-18)
  std::vector<float> uv;      // MeshData::uvs (2 floats each), empty: planar fallback
  std::vector<int32_t> uvidx; // MeshData::uvIndices, 3 per triangle
};
// A texture: ImageTexture texels (texel (x, y) at rgb[3 (y w + x)]) or a
// CheckerTexture's two colours and scale.
struct HTexture {
  uint32_t kind;  // TEX_IMAGE, TEX_CHECKER
  uint32_t w = 0, h = 0;
  std::vector<float> rgb;
  f3 c1, c2;
  float scale = 0.f;
};
// TriangleMesh's bounding box, as its constructor forms it: per-axis min and
// max over the vertices, then -/+ gl::epsilon (1e-3, constants.hpp:7).
static void mesh_box(const float *pos, uint32_t nv, f3 &lo, f3 &hi) {
  lo = s3(INFINITY);
  hi = s3(-INFINITY);
  for (uint32_t v = 0; v < nv; ++v) {
    const f3 p = mk3(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]);
    lo = mk3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
    hi = mk3(std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z));
  }
  lo = sub3(lo, s3(1e-3f));
  hi = add3(hi, s3(1e-3f));
}
struct HSphere { f3 c; float r; uint32_t mat, flags, seq; int light; };
struct HRect { int axis; float k, d0min, d0max, d1min, d1max; uint32_t mat, flags, seq; int light; };
struct HCurve { f3 cps[4]; float w0, w1; uint32_t mat, flags; };

struct pt_scene {
  std::vector<pt_material_desc> mats;
  std::vector<HTri> tris;
  std::vector<f3> normals3;  // three corner normals per smooth triangle
  std::vector<HMesh> meshes;
  std::vector<HSphere> spheres;
  std::vector<HRect> rects;
  std::vector<HCurve> curves;
  std::vector<HTexture> textures;
  uint32_t seq = 0;
  std::vector<float> env;
  uint32_t env_w = 0, env_h = 0;
  float env_scale = 1.f;
  bool env_dirty = true;      // the environment's device tables need building
  float env_integral = 0.f;   // of the environment distribution's function
  pt_camera_desc cam{};
  bool has_cam = false;
  bool committed = false;

  // device
  hipStream_t stream = nullptr;
  int device = -1;
  DevBuf d_nodes, d_qnodes, d_prims, d_prec, d_smooth, d_curve, d_mats, d_lights, d_env, d_env_func, d_env_cdf,
      d_env_mcdf, d_env_guide;
  DevBuf d_perm, d_perm_off;
  DevBuf d_tex, d_tex_desc, d_prim_tex, d_tri_uv, d_mesh_box;  // textures and texture coordinates
  DevBuf lb_in, lb_scratch, lb_nodes;  // device BVH build: input records, scratch, nodes before trimming
  PinnedBuf lb_host;                   // the input records, staged
  uint32_t perm_seed = 0xffffffffu;
  bool perm_valid = false;
  DevScene ds{};
  std::vector<DevLight> lights;
  pt_scene_info info{};
  // wavefront
  WaveStore ws;
  DevBuf d_off;
  pt_tuning tuning{};
  pt_build_cfg build_cfg{};
  float box_lo[3] = {0.f, 0.f, 0.f}, box_hi[3] = {0.f, 0.f, 0.f};  // all primitives
  // guided renders: processed mixtures per live entry, training records
  DevBuf g_mix, g_rec;
  uint32_t guide_tag = 0;  // tags of record sets, never reused
  DevBuf film, samples;
  uint32_t film_w = 0, film_h = 0;
  uint64_t film_spp = 0;
  size_t samples_n = 0;
  uint32_t n_cu = 256;
};

static bool g_has_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return false;
  return n > 0;
}

static pt_status ensure_stream(pt_scene *s) {
  if (s->stream) return PT_OK;
  if (!g_has_device()) return fail(PT_ERR_NO_DEVICE, "no HIP device visible");
  HIP_TRY(hipGetDevice(&s->device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, s->device));
  s->n_cu = (uint32_t)std::max(1, prop.multiProcessorCount);
  HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  return PT_OK;
}

// ------------------------------------------------------------------ C ABI: scene building
extern "C" const char *pt_last_error(void) { return g_err.c_str(); }
extern "C" const char *pt_version(void) { return "pathtracer_amd 0.1 (gfx950 wavefront)"; }

extern "C" pt_status pt_scene_create(pt_scene **out) {
  if (!out) return fail(PT_ERR_INVALID, "pt_scene_create: out is NULL");
  try {
    *out = new pt_scene();
  } catch (...) {
    return fail(PT_ERR_NOMEM, "pt_scene_create: out of host memory");
  }
  return PT_OK;
}

extern "C" void pt_scene_destroy(pt_scene *s) {
  if (!s) return;
  if (s->stream) {
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamDestroy(s->stream);
  }
  delete s;
}

static bool valid_kind(uint32_t k) { return k >= PT_MAT_LAMBERTIAN && k <= PT_MAT_KAJIYA_KAY; }

extern "C" pt_status pt_scene_add_material(pt_scene *s, const pt_material_desc *d, uint32_t *out_id) {
  if (!s || !d) return fail(PT_ERR_INVALID, "pt_scene_add_material: NULL argument");
  if (!valid_kind(d->kind)) return fail(PT_ERR_INVALID, "pt_scene_add_material: unknown material kind");
  if ((d->flags & ~0xffffff01u) != 0u) return fail(PT_ERR_INVALID, "pt_scene_add_material: unknown flags");
  if ((d->flags & 1u) && d->kind != PT_MAT_LAMBERTIAN && d->kind != PT_MAT_DISNEY)
    return fail(PT_ERR_UNSUPPORTED, "pt_scene_add_material: only Lambertian and Disney materials take a texture");
  if (s->mats.size() >= (1u << 24)) return fail(PT_ERR_INVALID, "too many materials");
  s->mats.push_back(*d);
  if (out_id) *out_id = (uint32_t)s->mats.size() - 1;
  s->committed = false;
  return PT_OK;
}

// f(i) for i in [0, n) on up to 16 host threads, in contiguous chunks.
template <class F>
static void parallel_for(size_t n, F f) {
  const size_t kMin = 1u << 16;
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  nt = (unsigned)std::min<size_t>(nt, (n + kMin - 1) / kMin);
  if (nt <= 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      const size_t b = t * per, e = std::min(n, b + per);
      for (size_t i = b; i < e; ++i) f(i);
    });
  for (auto &x : th) x.join();
}

static pt_status add_mesh(pt_scene *s, const char *fn, const float *pos, uint32_t nv, const int32_t *idx, uint32_t nt,
                          const float *nrm, uint32_t nn, const int32_t *nidx, uint32_t mat, uint32_t flags) {
  const std::string f(fn);
  if (!s || (!pos && nv) || (!idx && nt) || (nidx && !nrm && nt)) return fail(PT_ERR_INVALID, f + ": NULL argument");
  if (mat >= s->mats.size()) return fail(PT_ERR_INVALID, f + ": material id out of range");
  const size_t first = s->tris.size(), first_n = s->normals3.size();
  auto undo = [&](const char *msg) {
    s->tris.resize(first);
    s->normals3.resize(first_n);
    return fail(PT_ERR_INVALID, f + msg);
  };
  HMesh m{(uint32_t)first, nt, mat, flags, 0};
  s->tris.reserve(first + nt);
  for (uint32_t t = 0; t < nt; ++t) {
    HTri tr;
    for (int c = 0; c < 3; ++c) {
      int32_t v = idx[3 * t + c];
      if (v < 0 || (uint32_t)v >= nv) return undo(": vertex index out of range");
      tr.p[c] = mk3(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]);
    }
    tr.mesh = (uint32_t)s->meshes.size();
    tr.smooth = -1;
    if (nidx) {
      tr.smooth = (int)(s->normals3.size() / 3);
      for (int c = 0; c < 3; ++c) {
        int32_t v = nidx[3 * t + c];
        if (v < 0 || (uint32_t)v >= nn) return undo(": normal index out of range");
        s->normals3.push_back(mk3(nrm[3 * v], nrm[3 * v + 1], nrm[3 * v + 2]));
      }
    }
    s->tris.push_back(tr);
  }
  m.seq = s->seq++;
  m.nverts = nv;
  m.idx.assign(idx, idx + 3 * (size_t)nt);
  mesh_box(pos, nv, m.box_lo, m.box_hi);
  s->meshes.push_back(std::move(m));
  s->committed = false;
  return PT_OK;
}

extern "C" pt_status pt_scene_update_mesh(pt_scene *s, uint32_t mesh, const float *pos, uint32_t nv) {
  if (!s || (!pos && nv)) return fail(PT_ERR_INVALID, "pt_scene_update_mesh: NULL argument");
  if (mesh >= s->meshes.size()) return fail(PT_ERR_INVALID, "pt_scene_update_mesh: mesh index out of range");
  const HMesh &m = s->meshes[mesh];
  if (nv != m.nverts) return fail(PT_ERR_INVALID, "pt_scene_update_mesh: vertex count differs from the mesh's");
  for (size_t k = 0; k < 3 * (size_t)nv; ++k)
    if (!std::isfinite(pos[k])) return fail(PT_ERR_INVALID, "pt_scene_update_mesh: non-finite position");
  parallel_for(m.count, [&](size_t t) {
    HTri &tr = s->tris[m.first + t];
    for (int c = 0; c < 3; ++c) {
      const int32_t v = m.idx[3 * t + c];
      tr.p[c] = mk3(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]);
    }
  });
  mesh_box(pos, nv, s->meshes[mesh].box_lo, s->meshes[mesh].box_hi);
  s->committed = false;
  return PT_OK;
}

extern "C" pt_status pt_scene_set_mesh_uv(pt_scene *s, uint32_t mesh, const float *uv, uint32_t nuv,
                                          const int32_t *uvidx, uint32_t nt) {
  if (!s || (!uv && nuv) || (!uvidx && nt)) return fail(PT_ERR_INVALID, "pt_scene_set_mesh_uv: NULL argument");
  if (mesh >= s->meshes.size()) return fail(PT_ERR_INVALID, "pt_scene_set_mesh_uv: mesh index out of range");
  HMesh &m = s->meshes[mesh];
  if (nt != m.count) return fail(PT_ERR_INVALID, "pt_scene_set_mesh_uv: triangle count differs from the mesh's");
  for (size_t k = 0; k < 3 * (size_t)nt; ++k)
    if (uvidx[k] < 0 || (uint32_t)uvidx[k] >= nuv) return fail(PT_ERR_INVALID, "pt_scene_set_mesh_uv: uv index out of range");
  m.uv.assign(uv, uv + 2 * (size_t)nuv);
  m.uvidx.assign(uvidx, uvidx + 3 * (size_t)nt);
  s->committed = false;
  return PT_OK;
}

extern "C" pt_status pt_scene_add_texture_image(pt_scene *s, const float *rgb, uint32_t w, uint32_t h,
                                                uint32_t *out_id) {
  if (!s || !rgb) return fail(PT_ERR_INVALID, "pt_scene_add_texture_image: NULL argument");
  if (w < 2 || h < 2) return fail(PT_ERR_INVALID, "pt_scene_add_texture_image: width and height must be >= 2");
  if ((uint64_t)w * h >= (1ull << 28)) return fail(PT_ERR_INVALID, "pt_scene_add_texture_image: texture too large");
  HTexture t;
  t.kind = TEX_IMAGE;
  t.w = w;
  t.h = h;
  t.rgb.assign(rgb, rgb + 3 * (size_t)w * h);
  s->textures.push_back(std::move(t));
  if (out_id) *out_id = (uint32_t)s->textures.size() - 1;
  s->committed = false;
  return PT_OK;
}

extern "C" pt_status pt_scene_add_texture_checker(pt_scene *s, const float c1[3], const float c2[3], float scale,
                                                  uint32_t *out_id) {
  if (!s || !c1 || !c2) return fail(PT_ERR_INVALID, "pt_scene_add_texture_checker: NULL argument");
  HTexture t;
  t.kind = TEX_CHECKER;
  t.c1 = mk3(c1[0], c1[1], c1[2]);
  t.c2 = mk3(c2[0], c2[1], c2[2]);
  t.scale = scale;
  s->textures.push_back(std::move(t));
  if (out_id) *out_id = (uint32_t)s->textures.size() - 1;
  s->committed = false;
  return PT_OK;
}

extern "C" pt_status pt_scene_add_mesh(pt_scene *s, const float *pos, uint32_t nv, const int32_t *idx, uint32_t nt,
                                       uint32_t mat, uint32_t flags) {
  return add_mesh(s, "pt_scene_add_mesh", pos, nv, idx, nt, nullptr, 0, nullptr, mat, flags);
}

extern "C" pt_status pt_scene_add_mesh_normals(pt_scene *s, const float *pos, uint32_t nv, const int32_t *idx,
                                               uint32_t nt, const float *nrm, uint32_t nn, const int32_t *nidx,
                                               uint32_t mat, uint32_t flags) {
  if (!nidx && nt) return fail(PT_ERR_INVALID, "pt_scene_add_mesh_normals: NULL argument");
  return add_mesh(s, "pt_scene_add_mesh_normals", pos, nv, idx, nt, nrm, nn, nidx, mat, flags);
}

extern "C" pt_status pt_scene_add_sphere(pt_scene *s, const float c[3], float r, uint32_t mat, uint32_t flags) {
  if (!s || !c) return fail(PT_ERR_INVALID, "pt_scene_add_sphere: NULL argument");
  if (mat >= s->mats.size()) return fail(PT_ERR_INVALID, "pt_scene_add_sphere: material id out of range");
  if (!(r > 0.f) || !std::isfinite(r)) return fail(PT_ERR_INVALID, "pt_scene_add_sphere: radius must be > 0");
  s->spheres.push_back(HSphere{mk3(c[0], c[1], c[2]), r, mat, flags, s->seq++, -1});
  s->committed = false;
  return PT_OK;
}

extern "C" pt_status pt_scene_add_curve(pt_scene *s, const float cps[12], float width0, float width1,
                                        uint32_t type, uint32_t mat, uint32_t flags) {
  if (!s || !cps) return fail(PT_ERR_INVALID, "pt_scene_add_curve: NULL argument");
  if (type > PT_CURVE_RIBBON) return fail(PT_ERR_INVALID, "pt_scene_add_curve: unknown curve type");
  if (mat >= s->mats.size()) return fail(PT_ERR_INVALID, "pt_scene_add_curve: material id out of range");
  if (!(width0 >= 0.f) || !(width1 >= 0.f) || !std::isfinite(width0 + width1))
    return fail(PT_ERR_INVALID, "pt_scene_add_curve: widths must be finite and >= 0");
  HCurve c;
  for (int i = 0; i < 4; ++i) {
    c.cps[i] = mk3(cps[3 * i], cps[3 * i + 1], cps[3 * i + 2]);
    if (!std::isfinite(cps[3 * i] + cps[3 * i + 1] + cps[3 * i + 2]))
      return fail(PT_ERR_INVALID, "pt_scene_add_curve: non-finite control point");
  }
  c.w0 = width0;
  c.w1 = width1;
  c.mat = mat;
  c.flags = flags & (PT_PRIM_FLIP_FACE | PT_PRIM_FORCE_INSIDE);
  s->curves.push_back(c);
  s->committed = false;
  return PT_OK;
}

// evalBezier / evalBezierDerivative (primitives/bezier.hpp:12-35), the
// reference's operation order.
static f3 bezier_at(const f3 *c, float u) {
  float u1 = 1.0f - u;
  float b0 = u1 * u1 * u1, b1 = 3.0f * u * u1 * u1, b2 = 3.0f * u * u * u1, b3 = u * u * u;
  return add3(add3(add3(scl3(c[0], b0), scl3(c[1], b1)), scl3(c[2], b2)), scl3(c[3], b3));
}
static f3 bezier_deriv(const f3 *c, float u) {
  f3 d0 = scl3(sub3(c[1], c[0]), 3.0f), d1 = scl3(sub3(c[2], c[1]), 3.0f), d2 = scl3(sub3(c[3], c[2]), 3.0f);
  float u1 = 1.0f - u;
  float b0 = u1 * u1, b1 = 2.0f * u * u1, b2 = u * u;
  return add3(add3(scl3(d0, b0), scl3(d1, b1)), scl3(d2, b2));
}
// The 8 cylinder segments of a curve (curve.hpp:54-83): ends, axis, length,
// radius (the mean of the widths at the ends) and the end derivatives.
struct CurveSeg { f3 pa, axis, dpa, dpb; float len, r; };
static void curve_segments(const HCurve &c, CurveSeg seg[8]) {
  f3 P[9], dP[9];
  for (int i = 0; i <= 8; ++i) {
    float u = (float)i / 8.0f;
    P[i] = bezier_at(c.cps, u);
    dP[i] = bezier_deriv(c.cps, u);
  }
  for (int i = 0; i < 8; ++i) {
    float u0 = (float)i / 8.0f, u1 = (float)(i + 1) / 8.0f;
    float r0 = c.w0 + (c.w1 - c.w0) * u0, r1 = c.w0 + (c.w1 - c.w0) * u1;
    seg[i].pa = P[i];
    f3 dv = sub3(P[i + 1], P[i]);
    seg[i].axis = norm3(dv);
    seg[i].len = len3(dv);
    seg[i].r = 0.5f * (r0 + r1);
    seg[i].dpa = dP[i];
    seg[i].dpb = dP[i + 1];
  }
}

extern "C" pt_status pt_scene_add_rect(pt_scene *s, int axis, float k, float d0min, float d0max, float d1min,
                                       float d1max, uint32_t mat, uint32_t flags) {
  if (!s) return fail(PT_ERR_INVALID, "pt_scene_add_rect: NULL scene");
  if (axis < 0 || axis > 2) return fail(PT_ERR_INVALID, "pt_scene_add_rect: axis must be 0, 1 or 2");
  if (mat >= s->mats.size()) return fail(PT_ERR_INVALID, "pt_scene_add_rect: material id out of range");
  s->rects.push_back(HRect{axis, k, d0min, d0max, d1min, d1max, mat, flags, s->seq++, -1});
  s->committed = false;
  return PT_OK;
}

extern "C" pt_status pt_scene_set_env(pt_scene *s, const float *rgb, uint32_t w, uint32_t h, float scale) {
  if (!s || !rgb) return fail(PT_ERR_INVALID, "pt_scene_set_env: NULL argument");
  if (w < 2 || h < 2) return fail(PT_ERR_INVALID, "pt_scene_set_env: texture must be at least 2x2");
  s->env.assign(rgb, rgb + 3 * (size_t)w * h);
  s->env_w = w;
  s->env_h = h;
  s->env_scale = scale;
  s->env_dirty = true;
  s->committed = false;
  return PT_OK;
}

extern "C" pt_status pt_scene_set_medium(pt_scene *s, const float sigma_a[3], const float sigma_s[3],
                                         float sigma_scale, const float le[3], float le_scale, float g) {
  if (!s) return fail(PT_ERR_INVALID, "pt_scene_set_medium: NULL scene");
  if (!sigma_a) {
    s->ds.has_medium = 0;
    return PT_OK;
  }
  if (!sigma_s) return fail(PT_ERR_INVALID, "pt_scene_set_medium: sigma_s is NULL");
  float sa[3], ss[3];
  for (int c = 0; c < 3; ++c) {
    sa[c] = sigma_a[c] * sigma_scale;  // HomogeneousMedium::HomogeneousMedium (homogeneous.hpp:41-47)
    ss[c] = sigma_s[c] * sigma_scale;
    if (!(sa[c] >= 0.f) || !(ss[c] >= 0.f) || !std::isfinite(sa[c] + ss[c]))
      return fail(PT_ERR_INVALID, "pt_scene_set_medium: coefficients must be finite and >= 0");
  }
  if (!(sa[0] + ss[0] > 0.f || sa[1] + ss[1] > 0.f || sa[2] + ss[2] > 0.f))
    return fail(PT_ERR_INVALID, "pt_scene_set_medium: sigma_a + sigma_s must not be zero");
  if (!(g > -1.f && g < 1.f)) return fail(PT_ERR_INVALID, "pt_scene_set_medium: g must be in (-1, 1)");
  for (int c = 0; c < 3; ++c) {
    s->ds.med_sa[c] = sa[c];
    s->ds.med_ss[c] = ss[c];
  }
  (void)le;
  (void)le_scale;
  s->ds.med_g = g;
  s->ds.has_medium = 1;
  return PT_OK;
}

extern "C" pt_status pt_scene_set_camera(pt_scene *s, const pt_camera_desc *c) {
  if (!s || !c) return fail(PT_ERR_INVALID, "pt_scene_set_camera: NULL argument");
  if (!(c->fov > 0.f) || !(c->aspect > 0.f) || !(c->focal_dist > 0.f))
    return fail(PT_ERR_INVALID, "pt_scene_set_camera: fov, aspect and focal_dist must be > 0");
  s->cam = *c;
  s->has_cam = true;
  return PT_OK;
}

// ------------------------------------------------------------------ BVH build (binned SAH)
namespace {

struct BRef {
  float lo[3], hi[3];
  uint32_t prim;  // index into the unified primitive list
};

struct Builder {
  std::vector<BRef> refs;
  std::vector<float4> nodes;  // 4 per node
  std::vector<uint32_t> order;
  uint32_t leaf_size = 4, bins = 16;
  std::atomic<uint32_t> max_depth{0};
  std::atomic<uint32_t> node_count{0};

  static void bounds(const BRef *r, size_t n, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
    for (size_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], r[i].lo[a]);
        hi[a] = std::max(hi[a], r[i].hi[a]);
      }
  }
  static float area(const float lo[3], const float hi[3]) {
    float dx = std::max(0.f, hi[0] - lo[0]), dy = std::max(0.f, hi[1] - lo[1]), dz = std::max(0.f, hi[2] - lo[2]);
    return 2.f * (dx * dy + dy * dz + dz * dx);
  }
  static constexpr size_t kLeaf = ~(size_t)0;
  // Split [b, e): returns the split position (b < m < e), or kLeaf.
  size_t split(size_t b, size_t e, uint32_t depth) {
    size_t n = e - b;
    if (n <= 1) return kLeaf;
    BRef *r = refs.data();
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = b; i < e; ++i)
      for (int a = 0; a < 3; ++a) {
        float c = 0.5f * (r[i].lo[a] + r[i].hi[a]);
        clo[a] = std::min(clo[a], c);
        chi[a] = std::max(chi[a], c);
      }
    if (depth > 40) {  // bound the depth: median split on the widest centroid axis
      if (n <= leaf_size) return kLeaf;
      int ax = 0;
      for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
      size_t m = b + n / 2;
      std::nth_element(r + b, r + m, r + e, [ax](const BRef &x, const BRef &y) {
        return x.lo[ax] + x.hi[ax] < y.lo[ax] + y.hi[ax];
      });
      return m;
    }
    float plo[3], phi[3];
    bounds(r + b, n, plo, phi);
    float parea = area(plo, phi);
    const uint32_t NB = bins;
    float best_cost = INFINITY;
    int best_axis = -1;
    uint32_t best_bin = 0;
    for (int a = 0; a < 3; ++a) {
      float ext = chi[a] - clo[a];
      if (!(ext > 0.f)) continue;
      float scale = (float)NB / ext;
      uint32_t cnt[64] = {0};
      float blo[64][3], bhi[64][3];
      for (uint32_t k = 0; k < NB; ++k)
        for (int c = 0; c < 3; ++c) { blo[k][c] = INFINITY; bhi[k][c] = -INFINITY; }
      for (size_t i = b; i < e; ++i) {
        float c = 0.5f * (r[i].lo[a] + r[i].hi[a]);
        uint32_t k = std::min(NB - 1, (uint32_t)((c - clo[a]) * scale));
        ++cnt[k];
        for (int q = 0; q < 3; ++q) {
          blo[k][q] = std::min(blo[k][q], r[i].lo[q]);
          bhi[k][q] = std::max(bhi[k][q], r[i].hi[q]);
        }
      }
      float rarea[64];
      uint32_t rcnt[64];
      float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t acc = 0;
      for (int k = (int)NB - 1; k > 0; --k) {
        for (int q = 0; q < 3; ++q) { lo[q] = std::min(lo[q], blo[k][q]); hi[q] = std::max(hi[q], bhi[k][q]); }
        acc += cnt[k];
        rcnt[k] = acc;
        rarea[k] = area(lo, hi);
      }
      for (int q = 0; q < 3; ++q) { lo[q] = INFINITY; hi[q] = -INFINITY; }
      acc = 0;
      for (uint32_t k = 0; k + 1 < NB; ++k) {
        for (int q = 0; q < 3; ++q) { lo[q] = std::min(lo[q], blo[k][q]); hi[q] = std::max(hi[q], bhi[k][q]); }
        acc += cnt[k];
        if (acc == 0 || rcnt[k + 1] == 0) continue;
        float cost = area(lo, hi) * (float)acc + rarea[k + 1] * (float)rcnt[k + 1];
        if (cost < best_cost) { best_cost = cost; best_axis = a; best_bin = k; }
      }
    }
    // SAH: traversal cost 1, intersection cost 1 (relative)
    float leaf_cost = (float)n;
    float split_cost = (parea > 0.f) ? 1.f + best_cost / parea : INFINITY;
    if (n <= leaf_size && !(split_cost < leaf_cost)) return kLeaf;
    if (best_axis < 0) {
      if (n <= 16) return kLeaf;
      return b + n / 2;  // all centroids coincide: split by index
    }
    float ext = chi[best_axis] - clo[best_axis];
    float scale = (float)NB / ext;
    int ax = best_axis;
    float c0 = clo[ax];
    BRef *mid = std::partition(r + b, r + e, [&](const BRef &x) {
      float c = 0.5f * (x.lo[ax] + x.hi[ax]);
      return std::min(NB - 1, (uint32_t)((c - c0) * scale)) <= best_bin;
    });
    size_t m = (size_t)(mid - r);
    if (m == b || m == e) return b + n / 2;
    return m;
  }

  static void pad(float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
      float p = 1e-6f * (std::fabs(lo[a]) + std::fabs(hi[a])) + 1e-12f;
      lo[a] -= p;
      hi[a] += p;
    }
  }

  // Writes node `ni` for range [b, e) already known to be split at m.
  void write_node(uint32_t ni, size_t b, size_t m, size_t e, int c0, int c1) {
    float l0[3], h0[3], l1[3], h1[3];
    bounds(refs.data() + b, m - b, l0, h0);
    bounds(refs.data() + m, e - m, l1, h1);
    pad(l0, h0);
    pad(l1, h1);
    float4 *nd = &nodes[4 * (size_t)ni];
    nd[0] = make_float4(l0[0], h0[0], l0[1], h0[1]);
    nd[1] = make_float4(l1[0], h1[0], l1[1], h1[1]);
    nd[2] = make_float4(l0[2], h0[2], l1[2], h1[2]);
    int4 ch = make_int4(c0, c1, 0, 0);
    std::memcpy(&nd[3], &ch, sizeof(int4));
  }

  int build(size_t b, size_t e, uint32_t depth, bool parallel) {
    uint32_t md = max_depth.load(std::memory_order_relaxed);
    while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {
    }
    size_t m = split(b, e, depth);
    if (m == kLeaf) return leaf_code((uint32_t)b, (uint32_t)(e - b));
    uint32_t ni = node_count.fetch_add(1);
    int c0, c1;
    if (parallel && (e - b) > 200000) {
      int r0 = 0;
      std::thread th([&] { r0 = build(b, m, depth + 1, depth < 3); });
      c1 = build(m, e, depth + 1, depth < 3);
      th.join();
      c0 = r0;
    } else {
      c0 = build(b, m, depth + 1, false);
      c1 = build(m, e, depth + 1, false);
    }
    write_node(ni, b, m, e, c0, c1);
    return (int)ni;
  }
};

}  // namespace

// ------------------------------------------------------------------ BVH4 collapse
// Each 4-wide node takes the two children of a binary node and repeatedly
// opens its largest-area internal child until it has four children.  Node
// layout (8 x float4 = 128 bytes, one cache line): lo.x, hi.x, lo.y, hi.y,
// lo.z, hi.z of the four children (SoA), then the four child codes, then
// padding.  Parents precede children, so the root is node 0.
namespace {
struct Child {
  int code;
  float lo[3], hi[3];
};
float child_area(const Child &c) {
  float dx = c.hi[0] - c.lo[0], dy = c.hi[1] - c.lo[1], dz = c.hi[2] - c.lo[2];
  return dx * dy + dy * dz + dz * dx;
}
void bvh2_children(const std::vector<float4> &n2, int node, Child out[2]) {
  const float4 *nd = &n2[4 * (size_t)node];
  int4 ch;
  std::memcpy(&ch, &nd[3], sizeof(int4));
  out[0] = Child{ch.x, {nd[0].x, nd[0].z, nd[2].x}, {nd[0].y, nd[0].w, nd[2].y}};
  out[1] = Child{ch.y, {nd[1].x, nd[1].z, nd[2].z}, {nd[1].y, nd[1].w, nd[2].w}};
}
int collapse_node(const std::vector<float4> &n2, int node2, std::vector<float4> &n4, uint32_t depth,
                  uint32_t &max_depth) {
  if (depth > max_depth) max_depth = depth;
  Child ch[4];
  Child two[2];
  bvh2_children(n2, node2, two);
  int n = 2;
  ch[0] = two[0];
  ch[1] = two[1];
  while (n < 4) {
    int best = -1;
    float ba = -1.f;
    for (int i = 0; i < n; ++i)
      if (ch[i].code >= 0 && child_area(ch[i]) > ba) {
        ba = child_area(ch[i]);
        best = i;
      }
    if (best < 0) break;
    bvh2_children(n2, ch[best].code, two);
    ch[best] = two[0];
    ch[n++] = two[1];
  }
  int idx = (int)(n4.size() / 8);
  n4.resize(n4.size() + 8);
  int codes[4];
  for (int i = 0; i < 4; ++i) {
    if (i >= n) codes[i] = kEmptyChild;
    else if (ch[i].code >= 0) codes[i] = collapse_node(n2, ch[i].code, n4, depth + 1, max_depth);
    else codes[i] = ch[i].code;
  }
  float4 *nd = &n4[8 * (size_t)idx];
  float v[6][4];
  for (int i = 0; i < 4; ++i) {
    for (int a = 0; a < 3; ++a) {
      // empty slots get an inverted box; traversal also checks the code
      v[2 * a][i] = i < n ? ch[i].lo[a] : 1.f;
      v[2 * a + 1][i] = i < n ? ch[i].hi[a] : -1.f;
    }
  }
  for (int r = 0; r < 6; ++r) nd[r] = make_float4(v[r][0], v[r][1], v[r][2], v[r][3]);
  int4 cc = make_int4(codes[0], codes[1], codes[2], codes[3]);
  std::memcpy(&nd[6], &cc, sizeof(int4));
  nd[7] = make_float4(0.f, 0.f, 0.f, 0.f);
  return idx;
}
}  // namespace

static void collapse_bvh4(const std::vector<float4> &n2, std::vector<float4> &n4, uint32_t &max_depth) {
  n4.clear();
  n4.reserve(n2.size());
  max_depth = 0;
  collapse_node(n2, 0, n4, 1, max_depth);
}

// Renumbers 4-wide nodes breadth-first (the root stays node 0), so that the
// top levels are a prefix of the array; level_end[l] = nodes in levels 0..l.
static void bfs_renumber(std::vector<float4> &n4, std::vector<uint32_t> &level_end) {
  const size_t nn = n4.size() / 8;
  std::vector<int> q;
  std::vector<int> newid(nn, -1);
  q.reserve(nn);
  q.push_back(0);
  newid[0] = 0;
  level_end.clear();
  for (size_t lb = 0; lb < q.size();) {
    const size_t le = q.size();
    level_end.push_back((uint32_t)le);
    for (size_t k = lb; k < le; ++k) {
      int4 ch;
      std::memcpy(&ch, &n4[8 * (size_t)q[k] + 6], sizeof(int4));
      for (int c : {ch.x, ch.y, ch.z, ch.w})
        if (c >= 0) {
          newid[c] = (int)q.size();
          q.push_back(c);
        }
    }
    lb = le;
  }
  std::vector<float4> out(n4.size());
  for (size_t k = 0; k < q.size(); ++k) {
    const float4 *src = &n4[8 * (size_t)q[k]];
    float4 *dst = &out[8 * k];
    for (int r = 0; r < 8; ++r) dst[r] = src[r];
    int4 ch;
    std::memcpy(&ch, &dst[6], sizeof(int4));
    int *c = &ch.x;
    for (int j = 0; j < 4; ++j)
      if (c[j] >= 0) c[j] = newid[c[j]];
    std::memcpy(&dst[6], &ch, sizeof(int4));
  }
  n4.swap(out);
}

// Nodes staged in LDS for traversal: the largest prefix of whole levels
// among the first `levels` that fits PT_TOP_NODES.
static uint32_t top_node_count(const std::vector<uint32_t> &level_end, uint32_t levels) {
  uint32_t top = 0;
  for (size_t l = 0; l < level_end.size() && l < levels; ++l)
    if (level_end[l] <= PT_TOP_NODES) top = level_end[l];
  return top;
}

// ------------------------------------------------------------------ murmur / digit permutations
// gl::pbrt::murmurHash64A (utils/hash.hpp:11-56), for DigitPermutation tables.
static uint64_t murmur64a(const unsigned char *key, size_t len, uint64_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ull;
  const int r = 47;
  uint64_t h = seed ^ (len * m);
  const unsigned char *end = key + 8 * (len / 8);
  while (key != end) {
    uint64_t k;
    std::memcpy(&k, key, 8);
    key += 8;
    k *= m;
    k ^= k >> r;
    k *= m;
    h ^= k;
    h *= m;
  }
  switch (len & 7) {
    case 7: h ^= uint64_t(key[6]) << 48; [[fallthrough]];
    case 6: h ^= uint64_t(key[5]) << 40; [[fallthrough]];
    case 5: h ^= uint64_t(key[4]) << 32; [[fallthrough]];
    case 4: h ^= uint64_t(key[3]) << 24; [[fallthrough]];
    case 3: h ^= uint64_t(key[2]) << 16; [[fallthrough]];
    case 2: h ^= uint64_t(key[1]) << 8; [[fallthrough]];
    case 1: h ^= uint64_t(key[0]); h *= m;
  }
  h ^= h >> r;
  h *= m;
  h ^= h >> r;
  return h;
}

// 128-byte node bytes above which any-hit traversal reads the quantised nodes
// (one XCD's L2).
static const size_t kQanyBytes = (size_t)4 << 20;

static const int kHostPrimes[25] = {2,  3,  5,  7,  11, 13, 17, 19, 23, 29, 31, 37, 41,
                                    43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97};

// DigitPermutation (sampler/digitPerm.hpp:12-56) for all 25 bases.
static pt_status build_perm(uint32_t seed, DevBuf &d_tab, DevBuf &d_off) {
  std::vector<uint16_t> tab;
  std::vector<uint32_t> off(25);
  for (int bi = 0; bi < 25; ++bi) {
    int base = kHostPrimes[bi];
    off[bi] = (uint32_t)tab.size();
    int nd = 0;
    double inv = 1.0 / (double)base, ibm = 1.0;
    while (1.0 - (base - 1) * ibm < 1.0) { ++nd; ibm *= inv; }
    for (int di = 0; di < nd; ++di) {
      unsigned char buf[12];
      int32_t b32 = base, d32 = di;
      std::memcpy(buf, &b32, 4);
      std::memcpy(buf + 4, &d32, 4);
      std::memcpy(buf + 8, &seed, 4);
      uint64_t dseed = murmur64a(buf, 12, 0);
      for (int dv = 0; dv < base; ++dv)
        tab.push_back((uint16_t)permutation_element((uint32_t)dv, (uint32_t)base, (uint32_t)dseed));
    }
  }
  HIP_TRY(upload(d_tab, tab.data(), tab.size() * sizeof(uint16_t)));
  HIP_TRY(upload(d_off, off.data(), off.size() * sizeof(uint32_t)));
  return PT_OK;
}
static pt_status ensure_perm(pt_scene *s, uint32_t seed) {
  if (s->perm_valid && s->perm_seed == seed) return PT_OK;
  pt_status st = build_perm(seed, s->d_perm, s->d_perm_off);
  if (st != PT_OK) return st;
  s->perm_seed = seed;
  s->perm_valid = true;
  s->ds.perm = s->d_perm.as<uint16_t>();
  s->ds.perm_off = s->d_perm_off.as<uint32_t>();
  s->ds.perm_seed = seed;
  return PT_OK;
}

// ------------------------------------------------------------------ commit
// Search ranges for inverting the CDF cdf[0..n] with G buckets: for every
// u in [b/G, (b+1)/G) the answer max{i < n : cdf[i] <= u} lies in
// [out[b].x, out[b].y].  The bucket bounds are widened by 1e-6 relative, far
// more than the rounding of the device's bucket index u * G.
static void cdf_buckets(const float *cdf, uint32_t n, uint32_t G, uint2 *out) {
  auto answer = [&](double v) {  // cdf is non-decreasing and cdf[0] = 0 <= v
    const float *e = std::upper_bound(cdf, cdf + n, v, [](double a, float c) { return a < (double)c; });
    return (uint32_t)std::max<ptrdiff_t>(0, (e - cdf) - 1);
  };
  for (uint32_t b = 0; b < G; ++b) {
    const double lb = (double)b / G * (1.0 - 1e-6);
    const double ub = b + 1 == G ? 1.0 : (double)(b + 1) / G * (1.0 + 1e-6);
    out[b] = make_uint2(answer(lb), answer(ub));
  }
}

static float texel_max(const pt_scene *s, uint32_t x, uint32_t y) {
  const float *c = &s->env[3 * ((size_t)y * s->env_w + x)];
  return std::max(c[0], std::max(c[1], c[2]));
}

// -- textures: the texel pool and texture records, and the per-primitive
// data the texture coordinates need (only when a material is textured):
//   tex_desc  2 float4 per texture: image (TEX_IMAGE, w, h, first texel),
//             checker (TEX_CHECKER, scale, c1.x, c1.y) (c1.z, c2)
//   prim_tex  float4 per primitive id: triangle (mesh, uv record or ~0u);
//             rect (d0min, d0max, d1min, d1max)
//   tri_uv    4 float4 per triangle of a mesh with uvs: (p0, uv0.x)
//             (e1, uv0.y) (e2, uv1.x) (uv1.y, uv2, -)
//   mesh_box  2 float4 per mesh: TriangleMesh::meshAABB min, max
static pt_status commit_textures(pt_scene *s, DevScene &ds) {
  bool any = false;
  for (const pt_material_desc &m : s->mats)
    if (m.flags & 1u) {
      any = true;
      if ((m.flags >> 8) >= s->textures.size())
        return fail(PT_ERR_INVALID, "pt_scene_commit: a material names a texture that was not added");
    }
  ds.tex = nullptr;
  ds.tex_desc = ds.prim_tex = ds.tri_uv = ds.mesh_box = nullptr;
  if (!any) return PT_OK;
  ds.mat_kinds |= kMatTextured;
  auto bits = [](uint32_t u) { return __builtin_bit_cast(float, u); };
  std::vector<float4> desc;
  std::vector<float> texels;
  for (const HTexture &t : s->textures) {
    if (t.kind == TEX_IMAGE) {
      desc.push_back(make_float4(bits(TEX_IMAGE), bits(t.w), bits(t.h), bits((uint32_t)(texels.size() / 3))));
      desc.push_back(make_float4(0.f, 0.f, 0.f, 0.f));
      texels.insert(texels.end(), t.rgb.begin(), t.rgb.end());
    } else {
      desc.push_back(make_float4(bits(TEX_CHECKER), t.scale, t.c1.x, t.c1.y));
      desc.push_back(make_float4(t.c1.z, t.c2.x, t.c2.y, t.c2.z));
    }
  }
  if (texels.empty()) texels.assign(3, 0.f);
  const uint32_t ntri = (uint32_t)s->tris.size(), nsph = (uint32_t)s->spheres.size();
  const size_t nprim = ntri + nsph + s->rects.size() + 8 * s->curves.size();
  std::vector<float4> ptex(nprim, make_float4(0.f, 0.f, 0.f, 0.f)), box, tuv;
  for (const HMesh &m : s->meshes) {
    box.push_back(make_float4(m.box_lo.x, m.box_lo.y, m.box_lo.z, 0.f));
    box.push_back(make_float4(m.box_hi.x, m.box_hi.y, m.box_hi.z, 0.f));
  }
  for (uint32_t mi = 0; mi < (uint32_t)s->meshes.size(); ++mi) {
    const HMesh &m = s->meshes[mi];
    for (uint32_t t = 0; t < m.count; ++t) {
      const uint32_t id = m.first + t;
      uint32_t rec = 0xffffffffu;
      if (!m.uv.empty()) {
        rec = (uint32_t)(tuv.size() / 4);
        const HTri &tr = s->tris[id];
        const f3 e1 = sub3(tr.p[1], tr.p[0]), e2 = sub3(tr.p[2], tr.p[0]);
        const float *u0 = &m.uv[2 * (size_t)m.uvidx[3 * t]], *u1 = &m.uv[2 * (size_t)m.uvidx[3 * t + 1]],
                    *u2 = &m.uv[2 * (size_t)m.uvidx[3 * t + 2]];
        tuv.push_back(make_float4(tr.p[0].x, tr.p[0].y, tr.p[0].z, u0[0]));
        tuv.push_back(make_float4(e1.x, e1.y, e1.z, u0[1]));
        tuv.push_back(make_float4(e2.x, e2.y, e2.z, u1[0]));
        tuv.push_back(make_float4(u1[1], u2[0], u2[1], 0.f));
      }
      ptex[id] = make_float4(bits(mi), bits(rec), 0.f, 0.f);
    }
  }
  for (size_t i = 0; i < s->rects.size(); ++i) {
    const HRect &r = s->rects[i];
    ptex[ntri + nsph + i] = make_float4(r.d0min, r.d0max, r.d1min, r.d1max);
  }
  if (box.empty()) box.assign(2, make_float4(0.f, 0.f, 0.f, 0.f));
  if (tuv.empty()) tuv.assign(4, make_float4(0.f, 0.f, 0.f, 0.f));
  HIP_TRY(upload(s->d_tex, texels.data(), texels.size() * sizeof(float)));
  HIP_TRY(upload(s->d_tex_desc, desc.data(), desc.size() * sizeof(float4)));
  HIP_TRY(upload(s->d_prim_tex, ptex.data(), ptex.size() * sizeof(float4)));
  HIP_TRY(upload(s->d_tri_uv, tuv.data(), tuv.size() * sizeof(float4)));
  HIP_TRY(upload(s->d_mesh_box, box.data(), box.size() * sizeof(float4)));
  ds.tex = s->d_tex.as<float>();
  ds.tex_desc = s->d_tex_desc.as<float4>();
  ds.prim_tex = s->d_prim_tex.as<float4>();
  ds.tri_uv = s->d_tri_uv.as<float4>();
  ds.mesh_box = s->d_mesh_box.as<float4>();
  return PT_OK;
}

extern "C" pt_status pt_scene_commit(pt_scene *s, const pt_build_cfg *cfg) {
  if (!s) return fail(PT_ERR_INVALID, "pt_scene_commit: NULL scene");
  size_t nprim = s->tris.size() + s->spheres.size() + s->rects.size() + 8 * s->curves.size();
  if (nprim == 0) return fail(PT_ERR_INVALID, "pt_scene_commit: the scene has no objects");  // renderManager.hpp:64-68
  if (nprim >= (1u << 27)) return fail(PT_ERR_INVALID, "pt_scene_commit: too many primitives");
  pt_status st = ensure_stream(s);
  if (st != PT_OK) return st;
  auto t0 = std::chrono::steady_clock::now();

  const uint32_t ntri = (uint32_t)s->tris.size(), nsph = (uint32_t)s->spheres.size();
  // -- lights (insertion order; environment last)
  s->lights.clear();
  auto rect_light = [&](HRect &r) {
    r.light = -1;
    const pt_material_desc &md = s->mats[r.mat];
    if (md.kind != PT_MAT_EMITTER || !(md.p[3] > 0.f) || (r.flags & PT_PRIM_NOT_LIGHT)) return;
    f3 v[4];
    float k = r.k;
    if (r.axis == 0) {
      v[0] = mk3(k, r.d0min, r.d1min); v[1] = mk3(k, r.d0max, r.d1min); v[3] = mk3(k, r.d0min, r.d1max);
    } else if (r.axis == 1) {
      v[0] = mk3(r.d0min, k, r.d1min); v[1] = mk3(r.d0max, k, r.d1min); v[3] = mk3(r.d0min, k, r.d1max);
    } else {
      v[0] = mk3(r.d0min, r.d1min, k); v[1] = mk3(r.d0max, r.d1min, k); v[3] = mk3(r.d0min, r.d1max, k);
    }
    DevLight L{};
    L.type = LIGHT_QUAD;
    L.v0 = v[0];
    L.e1 = sub3(v[1], v[0]);
    L.e2 = sub3(v[3], v[0]);
    L.n = norm3(cross3(L.e1, L.e2));
    L.area = len3(cross3(L.e1, L.e2));
    L.le = scl3(mk3(md.p[0], md.p[1], md.p[2]), md.p[3]);
    L.axis = r.axis;
    L.k = r.k;
    L.d0min = r.d0min; L.d0max = r.d0max; L.d1min = r.d1min; L.d1max = r.d1max;
    r.light = (int)s->lights.size();
    s->lights.push_back(L);
  };
  auto sphere_light = [&](HSphere &sp) {
    sp.light = -1;
    const pt_material_desc &md = s->mats[sp.mat];
    if (md.kind != PT_MAT_EMITTER || !(md.p[3] > 0.f) || (sp.flags & PT_PRIM_NOT_LIGHT)) return;
    DevLight L{};
    L.type = LIGHT_SPHERE;
    L.c = sp.c;
    L.r = sp.r;
    L.area = (float)(4 * M_PI * sp.r * sp.r);
    L.le = scl3(mk3(md.p[0], md.p[1], md.p[2]), md.p[3]);
    sp.light = (int)s->lights.size();
    s->lights.push_back(L);
  };
  // rects and spheres each hold increasing sequence numbers: merge them
  for (size_t ri = 0, si = 0; ri < s->rects.size() || si < s->spheres.size();) {
    if (si >= s->spheres.size() || (ri < s->rects.size() && s->rects[ri].seq < s->spheres[si].seq))
      rect_light(s->rects[ri++]);
    else
      sphere_light(s->spheres[si++]);
  }
  if (!s->env.empty()) {
    DevLight L{};
    L.type = LIGHT_ENV;
    s->lights.push_back(L);
  }

  // -- primitive references and per-primitive info
  const bool device_build = cfg && (cfg->flags & PT_BUILD_DEVICE_LBVH);
  Builder B;
  B.leaf_size = (cfg && cfg->leaf_size) ? std::min<uint32_t>(cfg->leaf_size, 16u) : 4u;
  s->build_cfg = cfg ? *cfg : pt_build_cfg{};
  s->tuning.leaf_size = B.leaf_size;
  B.bins = (cfg && cfg->sah_bins) ? std::min<uint32_t>(std::max<uint32_t>(cfg->sah_bins, 2u), 64u) : 16u;
  B.refs.resize(nprim);  // the device build reads the primitive records instead
  std::vector<uint2> info(nprim);
  std::vector<float4> geom(nprim);
  // 5 float4 per triangle with shading normals, in triangle order
  std::vector<uint32_t> smooth_at(ntri);
  uint32_t n_smooth = 0;
  for (uint32_t i = 0; i < ntri; ++i) smooth_at[i] = s->tris[i].smooth >= 0 ? n_smooth++ : 0xffffffffu;
  std::vector<float4> smooth(5 * (size_t)n_smooth);
  parallel_for(ntri, [&](size_t i) {
    const HTri &t = s->tris[i];
    if (!device_build) {
      BRef &r = B.refs[i];
      for (int a = 0; a < 3; ++a) {
        r.lo[a] = std::min(comp3(t.p[0], a), std::min(comp3(t.p[1], a), comp3(t.p[2], a)));
        r.hi[a] = std::max(comp3(t.p[0], a), std::max(comp3(t.p[1], a), comp3(t.p[2], a)));
      }
      r.prim = (uint32_t)i;
    }
    const HMesh &m = s->meshes[t.mesh];
    f3 e1 = sub3(t.p[1], t.p[0]), e2 = sub3(t.p[2], t.p[0]);
    const uint32_t sm = smooth_at[i];
    if (sm != 0xffffffffu) {
      // p0, e1, e2 for the barycentrics of the hit, then the three corner normals
      const f3 *n = &s->normals3[3 * (size_t)t.smooth];
      float4 *q = &smooth[5 * (size_t)sm];
      q[0] = make_float4(t.p[0].x, t.p[0].y, t.p[0].z, n[0].x);
      q[1] = make_float4(e1.x, e1.y, e1.z, n[0].y);
      q[2] = make_float4(e2.x, e2.y, e2.z, n[0].z);
      q[3] = make_float4(n[1].x, n[1].y, n[1].z, n[2].x);
      q[4] = make_float4(n[2].y, n[2].z, 0.f, 0.f);
    }
    info[i] = make_uint2(m.mat | (PRIM_TRI << 24) | ((m.flags & 3u) << 28), sm);
    f3 ng = norm3(cross3(e1, e2));
    geom[i] = make_float4(ng.x, ng.y, ng.z, 0.f);
  });
  for (uint32_t i = 0; i < nsph; ++i) {
    const HSphere &sp = s->spheres[i];
    BRef &r = B.refs[ntri + i];
    r.lo[0] = sp.c.x - sp.r; r.lo[1] = sp.c.y - sp.r; r.lo[2] = sp.c.z - sp.r;
    r.hi[0] = sp.c.x + sp.r; r.hi[1] = sp.c.y + sp.r; r.hi[2] = sp.c.z + sp.r;
    r.prim = ntri + i;
    info[ntri + i] = make_uint2(sp.mat | (PRIM_SPHERE << 24) | ((sp.flags & 3u) << 28), (uint32_t)sp.light);
    geom[ntri + i] = make_float4(sp.c.x, sp.c.y, sp.c.z, sp.r);
  }
  for (uint32_t i = 0; i < s->rects.size(); ++i) {
    const HRect &rc = s->rects[i];
    uint32_t id = ntri + nsph + i;
    BRef &r = B.refs[id];
    int a = rc.axis, b0 = (a == 0) ? 1 : 0, b1 = (a == 2) ? 1 : 2;
    r.lo[a] = rc.k; r.hi[a] = rc.k;
    r.lo[b0] = rc.d0min; r.hi[b0] = rc.d0max;
    r.lo[b1] = rc.d1min; r.hi[b1] = rc.d1max;
    r.prim = id;
    info[id] = make_uint2(rc.mat | (PRIM_RECT << 24) | ((rc.flags & 3u) << 28), (uint32_t)rc.light);
    geom[id] = make_float4((float)a, rc.k, 0.f, 0.f);
  }
  // curves: 8 segments each, after the rects; their shading data per segment
  const uint32_t nrect = (uint32_t)s->rects.size();
  std::vector<float4> cgeom(4 * std::max<size_t>(1, 8 * s->curves.size()));
  std::vector<CurveSeg> csegs(8 * s->curves.size());
  for (uint32_t ci = 0; ci < (uint32_t)s->curves.size(); ++ci) {
    const HCurve &c = s->curves[ci];
    curve_segments(c, &csegs[8 * (size_t)ci]);
    for (uint32_t k = 0; k < 8; ++k) {
      const uint32_t si = 8 * ci + k, id = ntri + nsph + nrect + si;
      const CurveSeg &sg = csegs[si];
      f3 pb = add3(sg.pa, scl3(sg.axis, sg.len));
      BRef &r = B.refs[id];
      for (int a = 0; a < 3; ++a) {
        r.lo[a] = std::min(comp3(sg.pa, a), comp3(pb, a)) - sg.r;
        r.hi[a] = std::max(comp3(sg.pa, a), comp3(pb, a)) + sg.r;
      }
      r.prim = id;
      info[id] = make_uint2(c.mat | (PRIM_CURVE << 24) | ((c.flags & 3u) << 28), si);
      geom[id] = make_float4(0.f, 0.f, 0.f, 0.f);
      cgeom[4 * (size_t)si + 0] = make_float4(sg.pa.x, sg.pa.y, sg.pa.z, sg.len);
      cgeom[4 * (size_t)si + 1] = make_float4(sg.axis.x, sg.axis.y, sg.axis.z, sg.r);
      cgeom[4 * (size_t)si + 2] = make_float4(sg.dpa.x, sg.dpa.y, sg.dpa.z, 0.f);
      cgeom[4 * (size_t)si + 3] = make_float4(sg.dpb.x, sg.dpb.y, sg.dpb.z, 0.f);
    }
  }

  // -- primitive records (3 float4 each): triangle p0, e1, e2; sphere centre
  // and radius; rect k, extents, axis; curve segment start, axis, radius, length
  auto prim_record = [&](uint32_t id, float4 *P) {
    if (id < ntri) {
      const HTri &t = s->tris[id];
      f3 e1 = sub3(t.p[1], t.p[0]), e2 = sub3(t.p[2], t.p[0]);  // meshBVH.cpp:93-94
      uint32_t tag = id | (PRIM_TRI << 30);
      P[0] = make_float4(t.p[0].x, t.p[0].y, t.p[0].z, __builtin_bit_cast(float, tag));
      P[1] = make_float4(e1.x, e1.y, e1.z, 0.f);
      P[2] = make_float4(e2.x, e2.y, e2.z, 0.f);
    } else if (id < ntri + nsph) {
      const HSphere &sp = s->spheres[id - ntri];
      uint32_t tag = id | (PRIM_SPHERE << 30);
      P[0] = make_float4(sp.c.x, sp.c.y, sp.c.z, __builtin_bit_cast(float, tag));
      P[1] = make_float4(sp.r, 0.f, 0.f, 0.f);
      P[2] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else if (id < ntri + nsph + nrect) {
      const HRect &rc = s->rects[id - ntri - nsph];
      uint32_t tag = id | ((uint32_t)PRIM_RECT << 30);
      P[0] = make_float4(rc.k, rc.d0min, rc.d0max, __builtin_bit_cast(float, tag));
      P[1] = make_float4(rc.d1min, rc.d1max, (float)rc.axis, 0.f);
      P[2] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      const CurveSeg &sg = csegs[id - ntri - nsph - nrect];
      uint32_t tag = id | ((uint32_t)PRIM_CURVE << 30);
      P[0] = make_float4(sg.pa.x, sg.pa.y, sg.pa.z, __builtin_bit_cast(float, tag));
      P[1] = make_float4(sg.axis.x, sg.axis.y, sg.axis.z, sg.r);
      P[2] = make_float4(sg.len, 0.f, 0.f, 0.f);
    }
  };

  std::vector<float4> nodes4;
  std::vector<uint32_t> level_end;  // breadth-first node numbering: nodes in levels 0..l
  uint32_t depth4 = 0, n_nodes4 = 0;
  int root = 0;
  double bvh_ms = 0.0;
  if (device_build) {
    // -- BVH on the device: records in id order in, leaf order out
    const size_t rec_bytes = 3 * nprim * sizeof(float4);
    HIP_TRY(s->lb_host.alloc(rec_bytes));
    float4 *prims = s->lb_host.as<float4>();
    parallel_for(nprim, [&](size_t k) { prim_record((uint32_t)k, &prims[3 * k]); });
    HIP_TRY(s->lb_in.alloc(rec_bytes));
    HIP_TRY(hipMemcpyAsync(s->lb_in.p, prims, rec_bytes, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s->lb_scratch.alloc(lbvh_scratch_bytes((uint32_t)nprim)));
    HIP_TRY(s->lb_nodes.alloc(std::max<size_t>(1, nprim - 1) * 8 * sizeof(float4)));
    HIP_TRY(s->d_prims.alloc(rec_bytes));
    LbvhScratch sc{};
    lbvh_carve(s->lb_scratch.p, (uint32_t)nprim, sc);
    LbvhResult res{};
    auto b0 = std::chrono::steady_clock::now();
    HIP_TRY(lbvh_build(s->stream, s->lb_in.as<float4>(), (uint32_t)nprim, B.leaf_size, sc,
                       s->d_prims.as<float4>(), s->lb_nodes.as<float4>(), res));
    bvh_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - b0).count();
    depth4 = res.depth4;
    n_nodes4 = res.nodes4;
    root = res.root;
    level_end.assign(res.level_end, res.level_end + std::min<uint32_t>(res.depth4, 4u));
    for (int a = 0; a < 3; ++a) {
      s->box_lo[a] = res.lo[a];
      s->box_hi[a] = res.hi[a];
    }
    if (3 * depth4 + 3 > PT_STACK4)
      return fail(PT_ERR_UNSUPPORTED, "pt_scene_commit: BVH too deep for the traversal stack");
    HIP_TRY(s->d_nodes.alloc((size_t)n_nodes4 * 8 * sizeof(float4)));
    HIP_TRY(s->d_qnodes.alloc((size_t)n_nodes4 * PT_QNODE_F4 * sizeof(float4)));
    HIP_TRY(lbvh_quantize(s->stream, s->lb_nodes.as<float4>(), n_nodes4, s->d_qnodes.as<float4>()));
    HIP_TRY(hipMemcpyAsync(s->d_nodes.p, s->lb_nodes.p, (size_t)n_nodes4 * 8 * sizeof(float4),
                           hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
  } else {
    // -- BVH on the host: binned SAH, then collapsed into 4-wide nodes
    auto b0 = std::chrono::steady_clock::now();
    B.nodes.resize(4 * std::max<size_t>(1, 2 * nprim));
    Builder::bounds(B.refs.data(), nprim, s->box_lo, s->box_hi);
    root = B.build(0, nprim, 0, true);
    uint32_t nn = B.node_count.load();
    if (root < 0) {  // the whole scene is one leaf: wrap it in a root with both children on it
      uint32_t ni = nn++;
      B.write_node(ni, 0, nprim, nprim, root, root);
      // write_node bounded [0,nprim) and [nprim,nprim); give child 1 the same box
      B.nodes[4 * (size_t)ni + 1] = B.nodes[4 * (size_t)ni + 0];
      float4 z = B.nodes[4 * (size_t)ni + 2];
      B.nodes[4 * (size_t)ni + 2] = make_float4(z.x, z.y, z.x, z.y);
      // the root must be node 0: swap it into place
      if (ni != 0) return fail(PT_ERR_INVALID, "internal: leaf root not at node 0");
      root = (int)ni;
    }
    B.nodes.resize(4 * (size_t)nn);
    collapse_bvh4(B.nodes, nodes4, depth4);
    bfs_renumber(nodes4, level_end);
    bvh_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - b0).count();
    n_nodes4 = (uint32_t)(nodes4.size() / 8);
    if (3 * depth4 + 3 > PT_STACK4)
      return fail(PT_ERR_UNSUPPORTED, "pt_scene_commit: BVH too deep for the traversal stack");
    // primitives in leaf order
    std::vector<float4> prims(3 * nprim);
    parallel_for(nprim, [&](size_t k) { prim_record(B.refs[k].prim, &prims[3 * k]); });
    std::vector<float4> nodes_q((size_t)n_nodes4 * PT_QNODE_F4);
    parallel_for(n_nodes4, [&](size_t i) { quantize_node(&nodes4[8 * i], &nodes_q[PT_QNODE_F4 * i]); });
    HIP_TRY(upload(s->d_nodes, nodes4.data(), nodes4.size() * sizeof(float4)));
    HIP_TRY(upload(s->d_qnodes, nodes_q.data(), nodes_q.size() * sizeof(float4)));
    HIP_TRY(upload(s->d_prims, prims.data(), prims.size() * sizeof(float4)));
  }

  // -- environment tables (2-D piecewise-constant; same construction as the oracle)
  // (built again only after pt_scene_set_env)
  std::vector<float> efunc, ecdf, emarg, emcdf;
  std::vector<uint2> eguide;
  const bool env_build = !s->env.empty() && s->env_dirty;
  if (env_build) {
    uint32_t W = s->env_w, H = s->env_h;
    efunc.resize((size_t)W * H);
    ecdf.resize((size_t)(W + 1) * H);
    emarg.resize(H);
    emcdf.resize(H + 1);
    for (uint32_t y = 0; y < H; ++y) {
      uint32_t y0 = y < H - 2 ? y : H - 2;
      float st = sinf((float)M_PI * ((float)y + 0.5f) / (float)H);
      for (uint32_t x = 0; x < W; ++x) {
        uint32_t x0 = x < W - 2 ? x : W - 2;
        float v = std::max(std::max(texel_max(s, x0, y0), texel_max(s, x0 + 1, y0)),
                           std::max(texel_max(s, x0, y0 + 1), texel_max(s, x0 + 1, y0 + 1)));
        efunc[(size_t)y * W + x] = v * st;
      }
    }
    for (uint32_t y = 0; y < H; ++y) {
      float *cdf = &ecdf[(size_t)y * (W + 1)];
      const float *f = &efunc[(size_t)y * W];
      double acc = 0.0;
      cdf[0] = 0.f;
      for (uint32_t x = 0; x < W; ++x) {
        acc += (double)f[x] / W;
        cdf[x + 1] = (float)acc;
      }
      emarg[y] = (float)acc;
      if (acc > 0.0)
        for (uint32_t x = 1; x <= W; ++x) cdf[x] = (float)((double)cdf[x] / acc);
      else
        for (uint32_t x = 1; x <= W; ++x) cdf[x] = (float)x / (float)W;
    }
    double acc = 0.0;
    emcdf[0] = 0.f;
    for (uint32_t y = 0; y < H; ++y) {
      acc += (double)emarg[y] / H;
      emcdf[y + 1] = (float)acc;
    }
    s->env_integral = (float)acc;
    for (uint32_t y = 1; y <= H; ++y) emcdf[y] = acc > 0.0 ? (float)((double)emcdf[y] / acc) : (float)y / H;
    // search ranges: W buckets per row, then H for the marginal
    eguide.resize((size_t)H * W + H);
    for (uint32_t y = 0; y < H; ++y) cdf_buckets(&ecdf[(size_t)y * (W + 1)], W, W, &eguide[(size_t)y * W]);
    cdf_buckets(emcdf.data(), H, H, &eguide[(size_t)H * W]);
  }

  // -- upload
  {
    // one 32-byte record per primitive: geom, then info and padding
    std::vector<float4> rec(2 * geom.size());
    for (size_t i = 0; i < geom.size(); ++i) {
      rec[2 * i] = geom[i];
      uint32_t w[4] = {info[i].x, info[i].y, 0u, 0u};
      std::memcpy(&rec[2 * i + 1], w, sizeof(w));
    }
    HIP_TRY(upload(s->d_prec, rec.data(), rec.size() * sizeof(float4)));
  }
  HIP_TRY(upload(s->d_curve, cgeom.data(), cgeom.size() * sizeof(float4)));
  if (smooth.empty())
    HIP_TRY(s->d_smooth.alloc(sizeof(float4)));
  else
    HIP_TRY(upload(s->d_smooth, smooth.data(), smooth.size() * sizeof(float4)));
  HIP_TRY(upload(s->d_mats, s->mats.data(), s->mats.size() * sizeof(pt_material_desc)));
  if (s->lights.empty())
    HIP_TRY(s->d_lights.alloc(sizeof(DevLight)));
  else
    HIP_TRY(upload(s->d_lights, s->lights.data(), s->lights.size() * sizeof(DevLight)));
  if (env_build) {
    HIP_TRY(upload(s->d_env, s->env.data(), s->env.size() * sizeof(float)));
    HIP_TRY(upload(s->d_env_func, efunc.data(), efunc.size() * sizeof(float)));
    HIP_TRY(upload(s->d_env_cdf, ecdf.data(), ecdf.size() * sizeof(float)));
    HIP_TRY(upload(s->d_env_mcdf, emcdf.data(), emcdf.size() * sizeof(float)));
    HIP_TRY(upload(s->d_env_guide, eguide.data(), eguide.size() * sizeof(uint2)));
    s->env_dirty = false;
  }
  DevScene &ds = s->ds;
  ds.nodes = s->d_nodes.as<float4>();
  ds.qnodes = s->d_qnodes.as<float4>();
  ds.prims = s->d_prims.as<float4>();
  ds.prim_rec = s->d_prec.as<float4>();
  ds.tri_smooth = s->d_smooth.as<float4>();
  ds.curve_geom = s->d_curve.as<float4>();
  ds.mats = s->d_mats.as<pt_material_desc>();
  ds.lights = s->d_lights.as<DevLight>();
  ds.n_lights = (int)s->lights.size();
  ds.mat_kinds = 0u;
  for (const pt_material_desc &m : s->mats) ds.mat_kinds |= 1u << m.kind;
  st = commit_textures(s, ds);
  if (st != PT_OK) return st;
  ds.root = root;
  ds.top_nodes = (cfg && (cfg->flags & PT_BUILD_NO_LDS_NODES))
                     ? 0u
                     : top_node_count(level_end, (cfg && cfg->lds_levels) ? cfg->lds_levels : 3u);
  // Any-hit traversal reads the 64-byte quantised nodes once the 128-byte
  // nodes outgrow an XCD's 4 MiB L2: it is then bound by the node bytes it
  // fetches.  A smaller tree is served from LDS and L2, and the decode only
  // adds arithmetic (on a 100-node tree shadow rays took 16% longer).
  const uint32_t bflags = cfg ? cfg->flags : 0u;
  ds.qany = (bflags & PT_BUILD_ANY_HIT_QNODES)      ? 1u
            : (bflags & PT_BUILD_ANY_HIT_FULL_NODES) ? 0u
                                                     : (uint32_t)((size_t)n_nodes4 * 128u > kQanyBytes);
  ds.has_env = !s->env.empty();
  ds.env = ds.has_env ? s->d_env.as<float>() : nullptr;
  ds.env_w = (int)s->env_w;
  ds.env_h = (int)s->env_h;
  ds.env_scale = s->env_scale;
  ds.env_func = ds.has_env ? s->d_env_func.as<float>() : nullptr;
  ds.env_cdf = ds.has_env ? s->d_env_cdf.as<float>() : nullptr;
  ds.env_mcdf = ds.has_env ? s->d_env_mcdf.as<float>() : nullptr;
  ds.env_guide = ds.has_env ? s->d_env_guide.as<uint2>() : nullptr;
  ds.env_mguide = ds.has_env ? s->d_env_guide.as<uint2>() + (size_t)s->env_w * s->env_h : nullptr;
  ds.env_integral = s->env_integral;
  st = ensure_perm(s, s->perm_valid ? s->perm_seed : 0u);  // the tables do not depend on the geometry
  if (st != PT_OK) return st;

  auto t1 = std::chrono::steady_clock::now();
  pt_scene_info &in = s->info;
  in.n_materials = (uint32_t)s->mats.size();
  in.n_triangles = ntri;
  in.n_spheres = nsph;
  in.n_rects = (uint32_t)s->rects.size();
  in.n_lights = (uint32_t)s->lights.size();
  in.bvh_nodes = n_nodes4;
  in.bvh_depth = depth4;
  in.bvh_ms = bvh_ms;
  in.device_bytes = s->d_nodes.bytes + s->d_qnodes.bytes + s->d_prims.bytes + s->d_prec.bytes + s->d_smooth.bytes +
                    s->d_env.bytes +
                    s->d_env_func.bytes + s->d_env_cdf.bytes;
  in.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  in.flags = ds.qany ? PT_INFO_ANY_HIT_QNODES : 0u;
  s->committed = true;
  return PT_OK;
}

extern "C" pt_status pt_scene_get_info(const pt_scene *s, pt_scene_info *info) {
  if (!s || !info) return fail(PT_ERR_INVALID, "pt_scene_get_info: NULL argument");
  *info = s->info;
  return PT_OK;
}

// ------------------------------------------------------------------ wavefront buffers
// Layout of the wavefront of `cap` path slots: two live-path buffers, the
// radiance per slot and the shadow queue, each array 256-byte aligned.
// Partition p of a live-path buffer starts with at most R * ceil(n / (R P))
// paths (runs of R slots dealt round-robin) and never grows; its shadow rays,
// at most 4 per path, stay in partition p.
struct WaveLayout {
  uint64_t part_cap, shadow_part_cap;
  size_t lcap, scap, bytes;
};
static bool wave_layout(uint32_t cap, WaveLayout &l) {
  const size_t per_state = 20;   // 4-byte words per entry of one live-path buffer
  const size_t per_slot = 3;     // radiance per slot
  const size_t per_shadow = 11;  // 4-byte words per shadow-queue entry
  const size_t n_arrays = 2 * per_state + per_slot + per_shadow;
  const uint64_t run = PT_PART_RUN;
  l.part_cap = ((uint64_t)cap + run * PT_PARTS - 1) / (run * PT_PARTS) * run;
  l.shadow_part_cap = 4 * l.part_cap;
  if (l.shadow_part_cap * PT_PARTS >= (1ull << 32)) return false;
  l.lcap = (size_t)(l.part_cap * PT_PARTS);
  l.scap = (size_t)(l.shadow_part_cap * PT_PARTS);
  l.bytes = (2 * per_state * l.lcap + per_slot * (size_t)cap + per_shadow * l.scap) * 4 + n_arrays * 256;
  return true;
}
// Points the arrays of w into base (at least l.bytes), counters and stats
// into their own buffers.
static void carve_wave(char *base, const WaveLayout &l, uint32_t cap, uint32_t *counts, unsigned long long *stats,
                       Wave &w) {
  char *p = base;
  auto take = [&](size_t n) {
    char *r = p;
    p += ((n * 4 + 255) / 256) * 256;
    return r;
  };
  const size_t lcap = l.lcap, scap = l.scap;
  for (int b = 0; b < 2; ++b) {
    PathState &ps = w.st[b];
    ps.ox = (float *)take(lcap); ps.oy = (float *)take(lcap); ps.oz = (float *)take(lcap);
    ps.dx = (float *)take(lcap); ps.dy = (float *)take(lcap); ps.dz = (float *)take(lcap);
    ps.t = (float *)take(lcap); ps.prim = (int *)take(lcap);
    ps.bx = (float *)take(lcap); ps.by = (float *)take(lcap); ps.bz = (float *)take(lcap);
    ps.ppdf = (float *)take(lcap);
    ps.px = (float *)take(lcap); ps.py = (float *)take(lcap); ps.pz = (float *)take(lcap);
    ps.flags = (uint32_t *)take(lcap);
    ps.slot = (uint32_t *)take(lcap);
    ps.pbx = (float *)take(lcap); ps.pby = (float *)take(lcap); ps.pbz = (float *)take(lcap);
  }
  w.lx = (float *)take(cap); w.ly = (float *)take(cap); w.lz = (float *)take(cap);
  w.sox = (float *)take(scap); w.soy = (float *)take(scap); w.soz = (float *)take(scap);
  w.sdx = (float *)take(scap); w.sdy = (float *)take(scap); w.sdz = (float *)take(scap);
  w.stmax = (float *)take(scap);
  w.scr = (float *)take(scap); w.scg = (float *)take(scap); w.scb = (float *)take(scap);
  w.sslot = (uint32_t *)take(scap);
  w.counts = counts;
  w.stats = stats;
  w.cap = cap;
  w.part_cap = (uint32_t)l.part_cap;
  w.shadow_part_cap = (uint32_t)l.shadow_part_cap;
  w.sort_keys = 0u;
  w.rc = nullptr;
  w.rvis = nullptr;
}

static pt_status ensure_wave(WaveStore &ws, uint32_t cap) {
  if (ws.cap >= cap && ws.buf.p) return PT_OK;
  WaveLayout l;
  if (!wave_layout(cap, l)) return fail(PT_ERR_INVALID, "pt_render: wavefront too large");
  HIP_TRY(ws.buf.alloc(l.bytes));
  HIP_TRY(ws.counts.alloc(PT_COUNTER_WORDS * sizeof(uint32_t)));
  // the work queues start at zero (the kernels rewind them between launches)
  HIP_TRY(hipMemset(ws.counts.p, 0, PT_COUNTER_WORDS * sizeof(uint32_t)));
  HIP_TRY(ws.stats.alloc(CNT_N * sizeof(unsigned long long)));
  carve_wave(ws.buf.as<char>(), l, cap, ws.counts.as<uint32_t>(), ws.stats.as<unsigned long long>(), ws.wave);
  ws.res.release();
  ws.cap = cap;
  return PT_OK;
}

// The reservoir estimator's candidate buffers (4 candidates per slot).
static pt_status ensure_reservoir(WaveStore &ws) {
  Wave &w = ws.wave;
  if (w.rc) return PT_OK;
  const size_t cap = w.cap;
  const size_t rc_bytes = ((12 * cap * 4 + 255) / 256) * 256;
  HIP_TRY(ws.res.alloc(rc_bytes + 4 * cap));
  w.rc = ws.res.as<float>();
  w.rvis = reinterpret_cast<uint8_t *>(ws.res.as<char>() + rc_bytes);
  return PT_OK;
}

static DevCamera make_camera(const pt_camera_desc &c) {
  // PerspectiveCamera::PerspectiveCamera (camera.cpp:7-35)
  DevCamera d{};
  float theta = c.fov / 2.f;
  float hh = tanf(theta);
  float ph = 1.0f * hh * 2.0f;
  float pw = ph * c.aspect;
  f3 up = mk3(c.up[0], c.up[1], c.up[2]);
  f3 front = mk3(c.front[0], c.front[1], c.front[2]);
  d.lookat = neg3(norm3(front));
  d.h = norm3(cross3(up, d.lookat));
  d.v = norm3(cross3(d.lookat, d.h));
  d.horiz = scl3(scl3(d.h, pw), c.focal_dist);
  d.vert = scl3(scl3(d.v, ph), c.focal_dist);
  d.pos = mk3(c.pos[0], c.pos[1], c.pos[2]);
  d.pos1 = mk3(c.pos1[0], c.pos1[1], c.pos1[2]);
  d.focal_dist = c.focal_dist;
  d.radius = (c.aperture > 0.f && std::isfinite(c.aperture)) ? c.focal_dist / c.aperture * 0.5f : 0.f;
  d.shutter = c.shutter;
  d.normalize_dir = (int)c.normalize_dir;
  return d;
}

static void grid_factor(uint32_t n, uint32_t &a, uint32_t &b) {
  uint32_t x = 1;
  for (uint32_t f = 1; f * f <= n; ++f)
    if (n % f == 0) x = f;
  a = x;
  b = n / x;
}

// ------------------------------------------------------------------ render
// Checks a render configuration (the preconditions of SceneInfo::render,
// renderManager.hpp:51-70, and this renderer's limits).  spp: samples per
// pixel of the frame; sppc: samples this call renders per pixel.
static pt_status check_cfg(const char *who, const pt_scene *s, const pt_render_cfg *cfg, uint32_t &spp,
                           uint32_t &sppc) {
  const std::string w(who);
  if (!s || !cfg) return fail(PT_ERR_INVALID, w + ": NULL argument");
  if (!s->committed) return fail(PT_ERR_INVALID, w + ": scene not committed");
  if (!s->has_cam) return fail(PT_ERR_INVALID, w + ": camera is not initialized");  // renderManager.hpp:58-62
  if (cfg->width == 0 || cfg->height == 0 || cfg->spp_x == 0 || cfg->spp_y == 0)
    return fail(PT_ERR_INVALID, w + ": width, height, spp_x, spp_y must be > 0");
  if (cfg->estimator > PT_EST_VOLUME) return fail(PT_ERR_INVALID, w + ": unknown estimator");
  if (cfg->estimator == PT_EST_VOLUME && s->ds.has_env)
    return fail(PT_ERR_UNSUPPORTED, w + ": the volume estimator samples area lights only");
  if (cfg->sampler > PT_SAMPLER_STRATIFIED) return fail(PT_ERR_INVALID, w + ": unknown sampler");
  if (cfg->estimator == PT_EST_REFERENCE_NEE && cfg->light_samples > 4)
    return fail(PT_ERR_UNSUPPORTED, w + ": at most 4 light samples per vertex");
  spp = cfg->spp_x * cfg->spp_y;
  if (cfg->spp_begin >= spp) return fail(PT_ERR_INVALID, w + ": spp_begin out of range");
  sppc = cfg->spp_count ? std::min(cfg->spp_count, spp - cfg->spp_begin) : spp - cfg->spp_begin;
  const uint64_t npix = (uint64_t)cfg->width * cfg->height;
  if (npix * spp >= (1ull << 32)) return fail(PT_ERR_INVALID, w + ": width*height*spp must be < 2^32");
  return PT_OK;
}

static RenderParams make_params(const pt_render_cfg *cfg, uint32_t spp, uint32_t sppc) {
  RenderParams rp{};
  rp.width = cfg->width;
  rp.height = cfg->height;
  rp.spp_x = cfg->spp_x;
  rp.spp_y = cfg->spp_y;
  rp.spp = spp;
  rp.spp_begin = cfg->spp_begin;
  rp.spp_count = sppc;
  rp.max_depth = cfg->max_depth;
  rp.sampler = cfg->sampler;
  rp.estimator = cfg->estimator;
  rp.rr_depth = cfg->rr_depth;
  rp.light_samples = cfg->light_samples;
  grid_factor(std::max(1u, cfg->light_samples), rp.lx, rp.ly);
  rp.seed = cfg->seed;
  rp.bg[0] = cfg->bg[0]; rp.bg[1] = cfg->bg[1]; rp.bg[2] = cfg->bg[2];
  return rp;
}

// Per-frame jittered pixel offsets (FrameBuffer::setOffsets,
// framebuffer.cpp:131-145): off[k] and off[spp + k] for sample k.
static void pixel_offsets(const pt_render_cfg *cfg, uint32_t spp, float *off) {
  for (uint32_t i = 0; i < cfg->spp_x; ++i)
    for (uint32_t j = 0; j < cfg->spp_y; ++j) {
      uint32_t k = i * cfg->spp_y + j;
      float a0 = (float)i / (float)cfg->spp_x, a1 = (float)(i + 1) / (float)cfg->spp_x;
      float b0 = (float)j / (float)cfg->spp_y, b1 = (float)(j + 1) / (float)cfg->spp_y;
      off[k] = (float)(a0 + (a1 - a0) * (double)pt_rng(0xffffffffu - k, 0u, cfg->seed ^ 0x5bd1e995u));
      off[spp + k] = (float)(b0 + (b1 - b0) * (double)pt_rng(0xffffffffu - k, 1u, cfg->seed ^ 0x5bd1e995u));
    }
}

// Path vertices traced per path.  NEE+MIS and the reservoir estimator trace
// one vertex more: the hit that weights (MIS) or drops (reservoir) the last
// vertex's direct light.
static uint32_t n_vertices(const pt_render_cfg *cfg) {
  if (cfg->estimator == PT_EST_VOLUME) return 1u;  // single scattering: the camera ray only
  return (cfg->estimator == PT_EST_NEE_MIS || cfg->estimator == PT_EST_RESERVOIR) ? cfg->max_depth + 1
                                                                                  : cfg->max_depth;
}

// Launch grids.  Shading: at most shade_per_cu blocks of 256 per CU (4
// resident, each looping).  Traversal: one resident round (8 waves per SIMD
// = 8 blocks per CU), each wave refilling its lanes from its share of the
// rays.  Both are multiples of the partition count.
struct Grids {
  uint32_t shade, trav;
};
// Default shading grid cap, blocks per CU.  With the tile queue, 6, 8, 16 and
// 64 were within 1.2% of 32 on the 1.3M-triangle glass frame and the Disney +
// hair grid, and 8 lost 3% on the Cornell frame (fewer blocks than tiles).
#ifndef PT_SHADE_BLOCKS_PER_CU
#define PT_SHADE_BLOCKS_PER_CU 32u
#endif
static Grids make_grids(uint32_t n_cu, const pt_tuning &t) {
  const uint32_t tb = t.trav_blocks_per_cu ? t.trav_blocks_per_cu : 8u;
  const uint32_t sb = t.shade_blocks_per_cu ? t.shade_blocks_per_cu : PT_SHADE_BLOCKS_PER_CU;
  Grids g;
  g.shade = std::max<uint32_t>(PT_PARTS, n_cu * sb / PT_PARTS * PT_PARTS);
  g.trav = std::max<uint32_t>(PT_PARTS, n_cu * tb / PT_PARTS * PT_PARTS);
  return g;
}

// HIP-event timing of the stages (generate, extend, shade, shadow, film).
struct StageEvents {
  struct Ev { hipEvent_t a, b; int kind; };
  bool on = false;
  hipStream_t st = nullptr;
  std::vector<Ev> evs;
  int begin(int kind) {
    if (!on) return -1;
    Ev e;
    (void)hipEventCreate(&e.a);
    (void)hipEventCreate(&e.b);
    e.kind = kind;
    (void)hipEventRecord(e.a, st);
    evs.push_back(e);
    return (int)evs.size() - 1;
  }
  void end(int i) {
    if (i >= 0) (void)hipEventRecord(evs[i].b, st);
  }
  void add_to(pt_stats *stats) const {
    for (const Ev &ev : evs) {
      float t = 0.f;
      (void)hipEventElapsedTime(&t, ev.a, ev.b);
      double *dst[5] = {&stats->ms_generate, &stats->ms_extend, &stats->ms_shade, &stats->ms_shadow, &stats->ms_film};
      *dst[ev.kind] += t;
    }
  }
  ~StageEvents() {
    for (Ev &ev : evs) {
      (void)hipEventDestroy(ev.a);
      (void)hipEventDestroy(ev.b);
    }
  }
};

// The launches of one chunk: camera rays, nverts x (extend, shade, shadow),
// film.  b == nullptr renders rp's chunk of one scene; otherwise the batch's
// current chunk, and rp only carries the estimator and the largest chunk.
struct ChunkPlan {
  uint32_t max_paths;  // grid sizing: largest chunk
  uint32_t max_pix;
  uint32_t nverts;
  uint32_t estimator;
  uint32_t mat_kinds;
  bool count;
  Grids grids;
  // guided renders: the network and the material kinds it guides
  const GuideView *gv = nullptr;
  uint32_t gkinds = 0, infer_grid = 0, gverts = 0;
};
static void launch_chunk(hipStream_t st, const DevScene &s, const DevCamera &cam, const RenderParams &rp,
                         const Wave &w, const float *offx, const float *offy, float *film, float *samples,
                         const BatchView *b, const ChunkPlan &pl, StageEvents *ev) {
  const uint32_t n = pl.max_paths;
  int e = ev ? ev->begin(0) : -1;
  launch_generate(st, s, cam, rp, w, offx, offy, b);
  if (ev) ev->end(e);
  int qi = 0;
  for (uint32_t d = 0; d < pl.nverts; ++d) {
    e = ev ? ev->begin(1) : -1;
    launch_extend(st, s, w, qi, n, pl.grids.trav, pl.count, b);
    if (ev) ev->end(e);
    e = ev ? ev->begin(2) : -1;
    if (pl.gv && d < pl.gverts)
      launch_guide_infer(st, s, w, qi, *pl.gv, pl.gkinds, pl.infer_grid, d == 0u && w.gpix, rp.chunk_first,
                         rp.spp_count, n);
    if (pl.gv && d >= pl.gverts && (!w.grec || d > w.grec_verts)) {
      // a guided render past its guided vertices, with no training record
      // left to complete: the vertex is NEE+MIS with the path's last
      // mixture pdf (carried in the path state), which the plain kernel
      // shades with fewer registers and less code
      RenderParams rpm = rp;
      rpm.estimator = PT_EST_NEE_MIS;
      launch_shade(st, s, rpm, w, qi, d, n, pl.grids.shade, b, pl.mat_kinds);
    } else {
      launch_shade(st, s, rp, w, qi, d, n, pl.grids.shade, b, pl.mat_kinds);
    }
    if (ev) ev->end(e);
    e = ev ? ev->begin(3) : -1;
    if (pl.estimator != PT_EST_NAIVE && pl.estimator != PT_EST_ROULETTE)  // those trace no shadow rays
      launch_shadow(st, s, w, 4u * n, pl.grids.trav, pl.count, pl.estimator == PT_EST_RESERVOIR, b);
    if (ev) ev->end(e);
    qi ^= 1;
  }
  e = ev ? ev->begin(4) : -1;
  launch_film(st, rp, w, film, samples, w.stats, b, pl.max_pix);
  if (ev) ev->end(e);
  if (b) launch_next_chunk(st, *b);
}

static void fill_stats(pt_stats *stats, const unsigned long long *c, uint64_t paths, uint64_t launches, float ms) {
  std::memset(stats, 0, sizeof(*stats));
  stats->ms_total = ms;
  stats->paths = paths;
  stats->segments = c[CNT_SEGMENTS];
  stats->shadow_rays = c[CNT_SHADOW_RAYS];
  stats->node_visits = c[CNT_NODES];
  stats->prim_tests = c[CNT_PRIMS];
  stats->shadow_node_visits = c[CNT_SNODES];
  stats->shadow_prim_tests = c[CNT_SPRIMS];
  stats->nan_samples = c[CNT_NAN];
  stats->extend_launches = launches;
}

// Path slots in flight by default: 64M (about 22 GB of HBM).  A frame whose
// paths fit takes one chunk; on the 1.3M-triangle 1024^2 x 64-spp frame
// 8M / 16M / 32M / 64M slots took 122.7 / 105.4 / 95.4 / 92.2 ms: the late
// vertices of a chunk have few live paths, and a larger chunk keeps those
// launches full.
static const uint32_t kDefaultWavefront = 1u << 26;

extern "C" pt_status pt_scene_set_tuning(pt_scene *s, const pt_tuning *t) {
  if (!s || !t) return fail(PT_ERR_INVALID, "pt_scene_set_tuning: NULL argument");
  if (t->trav_blocks_per_cu > 64 || t->shade_blocks_per_cu > 256)
    return fail(PT_ERR_INVALID, "pt_scene_set_tuning: grid sizes out of range");
  if (t->sort_bits != 0 && t->sort_bits != 4)
    return fail(PT_ERR_INVALID, "pt_scene_set_tuning: sort_bits must be 0 or 4");
  s->tuning = *t;
  return PT_OK;
}

// One pass of a guided render: the guide, the kinds it guides, and whether
// the pass writes training records (one path in `stride`, `verts` vertices)
// and collects their samples.
struct GuidePass {
  pt_guide *g;
  uint32_t kinds;
  bool record;
  uint32_t stride, verts;
  uint32_t guided_verts;  // vertices 0 .. guided_verts-1 use the guide
  bool per_pixel;         // one query per pixel at the camera vertex
};

static pt_status render_frame(pt_scene *s, const pt_render_cfg *cfg, pt_stats *stats, const GuidePass *gp);

extern "C" pt_status pt_render(pt_scene *s, const pt_render_cfg *cfg, pt_stats *stats) {
  return render_frame(s, cfg, stats, nullptr);
}

static pt_status render_frame(pt_scene *s, const pt_render_cfg *cfg, pt_stats *stats, const GuidePass *gp) {
  uint32_t spp = 0, sppc = 0;
  pt_status st = check_cfg("pt_render", s, cfg, spp, sppc);
  if (st != PT_OK) return st;
  const uint64_t npix = (uint64_t)cfg->width * cfg->height;
  const uint64_t total = npix * sppc;
  st = ensure_stream(s);
  if (st != PT_OK) return st;
  if (cfg->sampler == PT_SAMPLER_HALTON_PERM) {
    st = ensure_perm(s, cfg->seed);
    if (st != PT_OK) return st;
  }
  uint32_t want = cfg->wavefront ? cfg->wavefront : (s->tuning.wavefront ? s->tuning.wavefront : kDefaultWavefront);
  uint64_t chunk_pix = std::max<uint64_t>(1, want / sppc);
  chunk_pix = std::min(chunk_pix, npix);
  uint32_t cap = (uint32_t)(chunk_pix * sppc);
  st = ensure_wave(s->ws, cap);
  if (st != PT_OK) return st;
  if (cfg->estimator == PT_EST_RESERVOIR) {
    st = ensure_reservoir(s->ws);
    if (st != PT_OK) return st;
  }

  // film
  if (s->film_w != cfg->width || s->film_h != cfg->height || !(cfg->flags & PT_RENDER_ACCUMULATE)) {
    HIP_TRY(s->film.alloc(npix * 3 * sizeof(float)));
    HIP_TRY(hipMemsetAsync(s->film.p, 0, npix * 3 * sizeof(float), s->stream));
    s->film_w = cfg->width;
    s->film_h = cfg->height;
    s->film_spp = 0;
  }
  float *samples = nullptr;
  if (cfg->flags & PT_RENDER_KEEP_SAMPLES) {
    s->samples_n = (size_t)total * 3;
    HIP_TRY(s->samples.alloc(s->samples_n * sizeof(float)));
    samples = s->samples.as<float>();
  }
  std::vector<float> off(2 * spp);
  pixel_offsets(cfg, spp, off.data());
  HIP_TRY(s->d_off.alloc(off.size() * sizeof(float)));
  HIP_TRY(hipMemcpyAsync(s->d_off.p, off.data(), off.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemsetAsync(s->ws.stats.p, 0, CNT_N * sizeof(unsigned long long), s->stream));

  RenderParams rp = make_params(cfg, spp, sppc);
  const DevCamera cam = make_camera(s->cam);
  ChunkPlan pl{};
  pl.nverts = n_vertices(cfg);
  pl.estimator = cfg->estimator;
  pl.mat_kinds = s->ds.mat_kinds;
  pl.count = (cfg->flags & PT_RENDER_COUNTERS) != 0;
  pl.grids = make_grids(s->n_cu, s->tuning);
  Wave wave = s->ws.wave;
  wave.sort_keys = ((cfg->flags & PT_RENDER_SORT_HITS) || s->tuning.sort_bits) ? 1u : 0u;
  GuideView gview{};
  uint32_t n_rec = 0;
  if (gp) {
    const size_t lcap = (size_t)wave.part_cap * PT_PARTS;
    HIP_TRY(s->g_mix.alloc(lcap * PT_G_MIX * sizeof(float)));
    wave.gmix = s->g_mix.as<float>();
    wave.grec = nullptr;
    wave.grec_stride = std::max(1u, gp->stride);
    wave.grec_verts = gp->verts;
    if (gp->record) {
      n_rec = (cap + wave.grec_stride - 1) / wave.grec_stride;
      const size_t bytes = (size_t)n_rec * gp->verts * PT_G_REC * sizeof(float);
      if (s->g_rec.bytes < bytes) {
        HIP_TRY(s->g_rec.alloc(bytes));
        HIP_TRY(hipMemset(s->g_rec.p, 0, s->g_rec.bytes));  // no stale tags
      }
      wave.grec = s->g_rec.as<float>();
    }
    gview = gp->g->view();
    pl.gv = &gview;
    pl.gkinds = gp->kinds;
    pl.gverts = gp->guided_verts;
    wave.gverts = gp->guided_verts;
    wave.gpix = gp->per_pixel ? 1u : 0u;
    wave.gkinds = gp->kinds;
    for (int a = 0; a < 3; ++a) {
      wave.glo[a] = gview.lo[a];
      wave.ginv[a] = gview.inv_ext[a];
    }
    pl.estimator = PT_EST_GUIDED;
    rp.estimator = PT_EST_GUIDED;
    // the network kernel holds 66 KB of LDS: two resident blocks per CU
    pl.infer_grid = std::max<uint32_t>(PT_PARTS, s->n_cu * 2u / PT_PARTS * PT_PARTS);
  }
  StageEvents ev;
  ev.on = (cfg->flags & PT_RENDER_TIMING) != 0;
  ev.st = s->stream;

  hipEvent_t tstart, tstop;
  HIP_TRY(hipEventCreate(&tstart));
  HIP_TRY(hipEventCreate(&tstop));
  HIP_TRY(hipEventRecord(tstart, s->stream));
  uint64_t launches = 0;
  for (uint64_t first = 0; first < total; first += cap) {
    uint32_t n = (uint32_t)std::min<uint64_t>(cap, total - first);
    rp.chunk_first = first;
    rp.chunk_paths = n;
    pl.max_paths = n;
    if (gp) wave.gpass = ++s->guide_tag;
    launch_chunk(s->stream, s->ds, cam, rp, wave, s->d_off.as<float>(), s->d_off.as<float>() + spp,
                 s->film.as<float>(), samples, nullptr, pl, &ev);
    if (gp && gp->record)
      launch_guide_collect(s->stream, wave, (n + wave.grec_stride - 1) / wave.grec_stride, wave.grec_verts,
                           wave.grec_stride, wave.gpass, gp->g->tset.as<float>(), gp->g->tcap,
                           gp->g->tcount.as<uint32_t>());
    launches += pl.nverts;
  }
  HIP_TRY(hipEventRecord(tstop, s->stream));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->film_spp += sppc;
  if (stats) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, tstart, tstop);
    unsigned long long c[CNT_N];
    HIP_TRY(hipMemcpy(c, s->ws.stats.p, sizeof(c), hipMemcpyDeviceToHost));
    fill_stats(stats, c, total, launches, ms);
    ev.add_to(stats);
  }
  (void)hipEventDestroy(tstart);
  (void)hipEventDestroy(tstop);
  return PT_OK;
}

// ------------------------------------------------------------------ render without allocation
// pt_render_noalloc keeps its wavefront (live paths, radiance, shadow queue,
// reservoir candidates), pixel offsets and counters in `scratch`, a file-scope
// device array that every GPU the library runs on holds until the process
// ends: kScratchBytes = 64 MiB, enough for kNoAllocSlots = 131,072 path slots
// and kNoAllocMaxSpp samples per pixel.  Its device address is looked up once
// per device (hipGetSymbolAddress) and kept in g_scratch; host code uses that
// address, never the array's name.  Calls on one GPU share the array: calls
// that overlap on one GPU (different threads or streams) race.
static constexpr size_t kScratchBytes = 64ull << 20;
static constexpr uint32_t kNoAllocSlots = 1u << 17;
static constexpr uint32_t kNoAllocMaxSpp = 1u << 16;
static constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static char *g_scratch[kMaxDevices];

// The device address of `scratch` on the current device (looked up on the
// first call on that device).
static pt_status scratch_address(int device, char **out) {
  if (device < 0 || device >= kMaxDevices) return fail(PT_ERR_UNSUPPORTED, "pt_render_noalloc: device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(PT_ERR_DEVICE, std::string("pt_render_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<char *>(p);
  }
  *out = g_scratch[device];
  return PT_OK;
}

extern "C" pt_status pt_render_noalloc(pt_scene *s, const pt_render_cfg *cfg, float *film, pt_stats *stats) {
  uint32_t spp = 0, sppc = 0;
  pt_status st = check_cfg("pt_render_noalloc", s, cfg, spp, sppc);
  if (st != PT_OK) return st;
  if (!film) return fail(PT_ERR_INVALID, "pt_render_noalloc: NULL film");
  if (cfg->flags & (PT_RENDER_KEEP_SAMPLES | PT_RENDER_TIMING | PT_RENDER_ACCUMULATE))
    return fail(PT_ERR_UNSUPPORTED,
                "pt_render_noalloc: PT_RENDER_KEEP_SAMPLES, PT_RENDER_TIMING and PT_RENDER_ACCUMULATE need allocations");
  if (cfg->wavefront > kNoAllocSlots)
    return fail(PT_ERR_INVALID, "pt_render_noalloc: wavefront above the maximum of 131072 path slots");
  if (spp > kNoAllocMaxSpp) return fail(PT_ERR_INVALID, "pt_render_noalloc: more than 65536 samples per pixel");
  if (cfg->sampler == PT_SAMPLER_HALTON_PERM && !(s->perm_valid && s->perm_seed == cfg->seed))
    return fail(PT_ERR_UNSUPPORTED, "pt_render_noalloc: the digit permutations of this seed are not built "
                                    "(a pt_render with the seed builds them)");
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != s->device) return fail(PT_ERR_INVALID, "pt_render_noalloc: call with the scene's device current");
  char *base = nullptr;
  st = scratch_address(dev, &base);
  if (st != PT_OK) return st;

  const uint64_t npix = (uint64_t)cfg->width * cfg->height;
  const uint64_t total = npix * sppc;
  const uint32_t want = cfg->wavefront ? cfg->wavefront : kNoAllocSlots;
  const uint64_t chunk_pix = std::min<uint64_t>(std::max<uint64_t>(1, want / sppc), npix);
  const uint64_t cap64 = chunk_pix * sppc;
  if (cap64 > kNoAllocSlots)
    return fail(PT_ERR_INVALID, "pt_render_noalloc: one pixel's samples exceed the 131072 path slots");
  const uint32_t cap = (uint32_t)cap64;

  // scratch layout: counters, stats, pixel offsets, reservoir candidates, wavefront
  WaveLayout l;
  if (!wave_layout(cap, l)) return fail(PT_ERR_INVALID, "pt_render_noalloc: wavefront too large");
  auto up = [](size_t n) { return (n + 255) / 256 * 256; };
  const size_t o_counts = 0;
  const size_t o_stats = o_counts + up(PT_COUNTER_WORDS * sizeof(uint32_t));
  const size_t o_off = o_stats + up(CNT_N * sizeof(unsigned long long));
  const size_t o_rc = o_off + up(2 * (size_t)spp * sizeof(float));
  const size_t rc_bytes = up(12 * (size_t)cap * sizeof(float));
  const size_t o_rvis = o_rc + rc_bytes;
  const size_t o_wave = o_rvis + up(4 * (size_t)cap);
  if (o_wave + l.bytes > kScratchBytes) return fail(PT_ERR_INVALID, "pt_render_noalloc: scratch too small");
  Wave wave{};
  carve_wave(base + o_wave, l, cap, reinterpret_cast<uint32_t *>(base + o_counts),
             reinterpret_cast<unsigned long long *>(base + o_stats), wave);
  if (cfg->estimator == PT_EST_RESERVOIR) {
    wave.rc = reinterpret_cast<float *>(base + o_rc);
    wave.rvis = reinterpret_cast<uint8_t *>(base + o_rvis);
  }
  wave.sort_keys = ((cfg->flags & PT_RENDER_SORT_HITS) || s->tuning.sort_bits) ? 1u : 0u;
  float *off = reinterpret_cast<float *>(base + o_off);

  const auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipMemsetAsync(film, 0, npix * 3 * sizeof(float), s->stream));
  HIP_TRY(hipMemsetAsync(wave.stats, 0, CNT_N * sizeof(unsigned long long), s->stream));
  // the work queues start at zero (the kernels rewind them between launches)
  HIP_TRY(hipMemsetAsync(wave.work(0, 0), 0, 3 * PT_PARTS * PT_COUNTER_STRIDE * sizeof(uint32_t), s->stream));
  launch_pixel_offsets(s->stream, cfg->spp_x, cfg->spp_y, cfg->seed, off);
  RenderParams rp = make_params(cfg, spp, sppc);
  const DevCamera cam = make_camera(s->cam);
  ChunkPlan pl{};
  pl.nverts = n_vertices(cfg);
  pl.estimator = cfg->estimator;
  pl.mat_kinds = s->ds.mat_kinds;
  pl.count = (cfg->flags & PT_RENDER_COUNTERS) != 0;
  pl.grids = make_grids(s->n_cu, s->tuning);
  uint64_t launches = 0;
  for (uint64_t first = 0; first < total; first += cap) {
    const uint32_t n = (uint32_t)std::min<uint64_t>(cap, total - first);
    rp.chunk_first = first;
    rp.chunk_paths = n;
    pl.max_paths = n;
    launch_chunk(s->stream, s->ds, cam, rp, wave, off, off + spp, film, nullptr, nullptr, pl, nullptr);
    launches += pl.nverts;
  }
  // per-pixel sums to means, as pt_film_read scales them (1 / spp in float)
  launch_film_scale(s->stream, film, npix * 3, 1.f / (float)sppc);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (stats) {
    unsigned long long c[CNT_N];
    HIP_TRY(hipMemcpy(c, wave.stats, sizeof(c), hipMemcpyDeviceToHost));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    fill_stats(stats, c, total, launches, (float)ms);
  }
  return PT_OK;
}

// ------------------------------------------------------------------ guided render
pt_status ptx_scene_box(pt_scene *s, float lo[3], float hi[3], int *device) {
  if (!s->committed) return fail(PT_ERR_INVALID, "pt_guide_create: scene not committed");
  pt_status st = ensure_stream(s);
  if (st != PT_OK) return st;
  for (int a = 0; a < 3; ++a) {
    lo[a] = s->box_lo[a];
    hi[a] = s->box_hi[a];
  }
  *device = s->device;
  return PT_OK;
}

extern "C" pt_status pt_render_guided(pt_scene *s, const pt_render_cfg *cfg, pt_guide *g, const pt_guide_cfg *gc,
                                      pt_stats *stats, pt_guide_stats *gs) {
  uint32_t spp = 0, sppc = 0;
  pt_status st = check_cfg("pt_render_guided", s, cfg, spp, sppc);
  if (st != PT_OK) return st;
  if (!g) return fail(PT_ERR_INVALID, "pt_render_guided: NULL guide");
  if (cfg->estimator != PT_EST_NEE_MIS)
    return fail(PT_ERR_UNSUPPORTED, "pt_render_guided: guiding extends the PT_EST_NEE_MIS estimator");
  st = ensure_stream(s);
  if (st != PT_OK) return st;
  if (g->device != s->device) return fail(PT_ERR_UNSUPPORTED, "pt_render_guided: guide and scene devices differ");
  const pt_guide_cfg c0{};
  const pt_guide_cfg &c = gc ? *gc : c0;
  const bool train = !(c.flags & PT_GUIDE_NO_TRAIN);
  // a frame rendered in one pass (no training) keeps its per-sample radiance
  if ((cfg->flags & PT_RENDER_KEEP_SAMPLES) && train)
    return fail(PT_ERR_UNSUPPORTED, "pt_render_guided: per-sample radiance is kept only with PT_GUIDE_NO_TRAIN");
  // Training schedule (measured on the slit-lit room at 1024^2, 256 spp, three
  // films each): passes of 4, 8 and 4 spp (a sixteenth of the frame) with 96
  // Adam steps after each gave the error of passes of 1, 2, ... 16, 1 spp (an
  // eighth) with 64 steps each (relMSE 0.389 against 0.383) in 526 ms instead
  // of 556: fewer, larger passes keep the wavefront full and sync the host less.
  const uint32_t train_spp = train ? std::min(sppc, c.train_spp ? c.train_spp : std::max(1u, sppc / 16u)) : 0u;
  const uint32_t batch = c.batch ? std::min(c.batch, 1u << 20) : 16384u;
  const uint32_t max_steps = c.max_steps ? c.max_steps : 96u;
  const float lr = c.lr > 0.f ? c.lr : 1e-3f;
  GuidePass gp{};
  gp.g = g;
  gp.kinds = c.kinds ? c.kinds
                     : ((1u << PT_MAT_LAMBERTIAN) | (1u << PT_MAT_CONDUCTOR) | (1u << PT_MAT_DISNEY));
  gp.stride = c.record_stride ? c.record_stride : 1u;

  // Guiding deeper vertices compounds the throughput growth the defensive
  // mixture allows (up to 1 / (1 - c) per vertex): measured on the slit-lit
  // room at 256^2, 256 spp, guiding every vertex raised the relMSE 8x over the
  // plain tracer.  Guiding the first vertex alone lowers it as much as guiding
  // the first two (trimmed relMSE 0.258 and 0.259 against 0.262 and 0.254 over
  // two runs of three films each; plain 0.305) for half the network queries.
  gp.guided_verts = c.guide_verts ? c.guide_verts : 1u;
  // The camera vertices of a pixel lie within a pixel's footprint of one
  // another, where the network's input (position and normal, 32 bins per
  // axis) barely changes: they share one query, at the pixel's first sample
  // in the chunk.  Each path still decides guidability by its own material,
  // and samples and weights with the mixture it uses, so the estimate stays
  // unbiased; the training records hold each path's own inputs.
  gp.per_pixel = !(c.flags & PT_GUIDE_PATH_QUERIES);
  // records are written at guided vertices only: by default one per guided vertex
  gp.verts = std::min(c.record_verts ? c.record_verts : gp.guided_verts, 64u);
  if (train) {
    g->tcap = 1u << 21;
    HIP_TRY(g->tset.alloc((size_t)g->tcap * PT_G_SAMPLE * sizeof(float)));
    HIP_TRY(g->tcount.alloc(256));
  }
  pt_stats acc{};
  pt_guide_stats gacc{};
  double ms_train = 0.0;
  hipEvent_t ta, tb;
  HIP_TRY(hipEventCreate(&ta));
  HIP_TRY(hipEventCreate(&tb));
  auto add = [&](const pt_stats &p) {
    acc.paths += p.paths;
    acc.segments += p.segments;
    acc.shadow_rays += p.shadow_rays;
    acc.node_visits += p.node_visits;
    acc.prim_tests += p.prim_tests;
    acc.shadow_node_visits += p.shadow_node_visits;
    acc.shadow_prim_tests += p.shadow_prim_tests;
    acc.nan_samples += p.nan_samples;
    acc.extend_launches += p.extend_launches;
    acc.ms_total += p.ms_total;
    acc.ms_generate += p.ms_generate;
    acc.ms_extend += p.ms_extend;
    acc.ms_shade += p.ms_shade;
    acc.ms_shadow += p.ms_shadow;
    acc.ms_film += p.ms_film;
  };
  uint32_t begin = cfg->spp_begin, done = 0, pass_spp = c.first_pass_spp ? c.first_pass_spp : 4u;
  bool first = true;
  pt_status result = PT_OK;
  while (done < sppc && result == PT_OK) {
    const bool training = done < train_spp;
    const uint32_t cnt = training ? std::min(pass_spp, train_spp - done) : sppc - done;
    pt_render_cfg pc = *cfg;
    pc.spp_begin = begin;
    pc.spp_count = cnt;
    if (!first) pc.flags |= PT_RENDER_ACCUMULATE;
    gp.record = training;
    if (training) {
      if (hipMemsetAsync(g->tcount.p, 0, sizeof(uint32_t), s->stream) != hipSuccess) {
        result = fail(PT_ERR_DEVICE, "pt_render_guided: memset failed");
        break;
      }
    }
    pt_stats ps{};
    result = render_frame(s, &pc, &ps, &gp);
    if (result != PT_OK) break;
    add(ps);
    gacc.passes += 1;
    if (training) {
      uint32_t n = 0;
      if (hipMemcpy(&n, g->tcount.p, sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) {
        result = fail(PT_ERR_DEVICE, "pt_render_guided: reading the sample count failed");
        break;
      }
      n = std::min(n, g->tcap);
      gacc.samples += n;
      if (n > 0) {
        // a fixed number of steps per pass: the early passes' few samples are revisited
        const uint32_t bsz = std::min(batch, n);
        const uint32_t steps = max_steps;
        (void)hipEventRecord(ta, s->stream);
        for (uint32_t k = 0; k < steps && result == PT_OK; ++k) {
          float *lp = (k + 1 == steps) ? &gacc.last_loss : nullptr;
          result = ptx_guide_step(g, s->stream, g->tset.as<float>(), n, bsz, 0x9E3779B9u * (s->guide_tag + 1u) + k + 1u,
                                  lr, lp);
          gacc.steps += 1;
        }
        (void)hipEventRecord(tb, s->stream);
        (void)hipEventSynchronize(tb);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ta, tb);
        ms_train += ms;
      }
      pass_spp *= 2u;
    }
    begin += cnt;
    done += cnt;
    first = false;
  }
  (void)hipEventDestroy(ta);
  (void)hipEventDestroy(tb);
  if (result != PT_OK) return result;
  acc.ms_total += ms_train;
  gacc.ms_train = ms_train;
  if (stats) *stats = acc;
  if (gs) *gs = gacc;
  return PT_OK;
}


extern "C" pt_status pt_film_read(pt_scene *s, float *out) {
  if (!s || !out) return fail(PT_ERR_INVALID, "pt_film_read: NULL argument");
  if (!s->film.p || s->film_spp == 0) return fail(PT_ERR_INVALID, "pt_film_read: nothing rendered");
  size_t n = (size_t)s->film_w * s->film_h * 3;
  HIP_TRY(hipMemcpy(out, s->film.p, n * sizeof(float), hipMemcpyDeviceToHost));
  float inv = 1.f / (float)s->film_spp;
  for (size_t i = 0; i < n; ++i) out[i] *= inv;  // FrameBuffer average (renderManager.hpp:146-150)
  return PT_OK;
}

extern "C" pt_status pt_film_read_samples(pt_scene *s, float *out, size_t n) {
  if (!s || !out) return fail(PT_ERR_INVALID, "pt_film_read_samples: NULL argument");
  if (!s->samples.p || n != s->samples_n) return fail(PT_ERR_INVALID, "pt_film_read_samples: size mismatch");
  HIP_TRY(hipMemcpy(out, s->samples.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return PT_OK;
}

// ------------------------------------------------------------------ batches
// One render of a batch: its scene, configuration and film.
struct BatchInst {
  pt_scene *scene = nullptr;
  pt_render_cfg cfg{};
  DevBuf film, perm, perm_off;
  uint32_t perm_seed = 0xffffffffu;
  uint32_t film_w = 0, film_h = 0;
  uint64_t film_spp = 0;
};

struct pt_batch {
  std::vector<std::unique_ptr<BatchInst>> insts;
  hipStream_t stream = nullptr;
  int device = -1;
  uint32_t n_cu = 256;
  pt_tuning tuning{};
  WaveStore ws;
  // per-render tables and the chunk schedule (device copies of the host vectors)
  DevBuf d_scenes, d_params, d_cams, d_offs, d_off_first, d_films, d_segs, d_chunks, d_cur, d_slots;
  // the captured launches of one chunk, per shading material set (0: the
  // basic set, 1: every kind), and what they were captured for
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<unsigned char> graph_key[2];
  void drop_graph() {
    for (int k = 0; k < 2; ++k) {
      if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
      if (graph[k]) (void)hipGraphDestroy(graph[k]);
      exec[k] = nullptr;
      graph[k] = nullptr;
      graph_key[k].clear();
    }
  }
};

extern "C" pt_status pt_batch_create(pt_batch **out) {
  if (!out) return fail(PT_ERR_INVALID, "pt_batch_create: out is NULL");
  try {
    *out = new pt_batch();
  } catch (...) {
    return fail(PT_ERR_NOMEM, "pt_batch_create: out of host memory");
  }
  return PT_OK;
}

extern "C" void pt_batch_destroy(pt_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->drop_graph();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

static pt_status batch_check(const char *who, pt_scene *scene, const pt_render_cfg *cfg) {
  uint32_t spp, sppc;
  pt_status st = check_cfg(who, scene, cfg, spp, sppc);
  if (st != PT_OK) return st;
  if (cfg->flags & PT_RENDER_KEEP_SAMPLES)
    return fail(PT_ERR_UNSUPPORTED, std::string(who) + ": batches do not keep per-sample radiance");
  return PT_OK;
}

extern "C" pt_status pt_batch_add(pt_batch *b, pt_scene *scene, const pt_render_cfg *cfg, uint32_t *out_index) {
  if (!b) return fail(PT_ERR_INVALID, "pt_batch_add: NULL batch");
  pt_status st = batch_check("pt_batch_add", scene, cfg);
  if (st != PT_OK) return st;
  if (b->insts.size() >= (1u << 20)) return fail(PT_ERR_INVALID, "pt_batch_add: too many renders");
  auto in = std::make_unique<BatchInst>();
  in->scene = scene;
  in->cfg = *cfg;
  b->insts.push_back(std::move(in));
  if (out_index) *out_index = (uint32_t)(b->insts.size() - 1);
  return PT_OK;
}

extern "C" pt_status pt_batch_set_cfg(pt_batch *b, uint32_t index, const pt_render_cfg *cfg) {
  if (!b) return fail(PT_ERR_INVALID, "pt_batch_set_cfg: NULL batch");
  if (index >= b->insts.size()) return fail(PT_ERR_INVALID, "pt_batch_set_cfg: index out of range");
  pt_status st = batch_check("pt_batch_set_cfg", b->insts[index]->scene, cfg);
  if (st != PT_OK) return st;
  b->insts[index]->cfg = *cfg;
  return PT_OK;
}

extern "C" pt_status pt_batch_set_tuning(pt_batch *b, const pt_tuning *t) {
  if (!b || !t) return fail(PT_ERR_INVALID, "pt_batch_set_tuning: NULL argument");
  if (t->trav_blocks_per_cu > 64 || t->shade_blocks_per_cu > 256)
    return fail(PT_ERR_INVALID, "pt_batch_set_tuning: grid sizes out of range");
  if (t->sort_bits != 0 && t->sort_bits != 4)
    return fail(PT_ERR_INVALID, "pt_batch_set_tuning: sort_bits must be 0 or 4");
  b->tuning = *t;
  return PT_OK;
}

// Cuts the concatenated path space of the renders (each pixel-major, whole
// pixels, taken in `order`) into chunks of at most cap slots, and each chunk
// into one segment per render it touches.
static void plan_chunks(const std::vector<uint32_t> &npix, const std::vector<uint32_t> &sppc,
                        const std::vector<uint32_t> &order, uint32_t cap, std::vector<ChunkSeg> &segs,
                        std::vector<ChunkInfo> &chunks) {
  segs.clear();
  chunks.clear();
  ChunkInfo cur{0u, 0u, 0u, 0u};
  for (uint32_t i : order) {
    uint32_t done = 0;
    while (done < npix[i]) {
      uint32_t room = (cap - cur.paths) / sppc[i];
      if (room == 0) {  // the chunk cannot take a whole pixel of this render
        chunks.push_back(cur);
        cur = ChunkInfo{(uint32_t)segs.size(), 0u, 0u, 0u};
        continue;
      }
      uint32_t m = std::min(room, npix[i] - done);
      ChunkSeg sg{i, cur.paths, cur.npix, done, m};
      segs.push_back(sg);
      cur.n_segs += 1;
      cur.paths += m * sppc[i];
      cur.npix += m;
      done += m;
    }
  }
  if (cur.n_segs) chunks.push_back(cur);
}

template <class T>
static hipError_t upload_vec(DevBuf &d, const std::vector<T> &v) {
  return upload(d, v.data(), std::max<size_t>(1, v.size()) * sizeof(T));
}

static pt_status batch_render(pt_batch *b, const pt_batch_cfg *bc, pt_stats *stats, uint32_t spp_div) {
  const size_t ni = b->insts.size();
  if (ni == 0) return fail(PT_ERR_INVALID, "pt_batch_render: the batch is empty");
  const uint32_t flags = bc ? bc->flags : 0u;
  const uint32_t est = b->insts[0]->cfg.estimator;
  std::vector<uint32_t> spp(ni), sppc(ni), npix(ni);
  uint32_t nverts = 0, mat_kinds = 0, max_sppc = 0;
  uint64_t total = 0;
  for (size_t i = 0; i < ni; ++i) {
    BatchInst &in = *b->insts[i];
    pt_status st = batch_check("pt_batch_render", in.scene, &in.cfg);
    if (st != PT_OK) return st;
    if (in.cfg.estimator != est)
      return fail(PT_ERR_UNSUPPORTED, "pt_batch_render: all renders of a batch must use one estimator");
    st = ensure_stream(in.scene);
    if (st != PT_OK) return st;
    if (b->device < 0) b->device = in.scene->device;
    if (in.scene->device != b->device)
      return fail(PT_ERR_UNSUPPORTED, "pt_batch_render: the scenes of a batch must be on one device");
    spp[i] = in.cfg.spp_x * in.cfg.spp_y;
    const uint32_t left = spp[i] - in.cfg.spp_begin;
    sppc[i] = in.cfg.spp_count ? std::min(in.cfg.spp_count, left) : left;
    if (spp_div > 1) sppc[i] = std::max(1u, sppc[i] / spp_div);
    npix[i] = in.cfg.width * in.cfg.height;
    nverts = std::max(nverts, n_vertices(&in.cfg));
    mat_kinds |= in.scene->ds.mat_kinds;
    max_sppc = std::max(max_sppc, sppc[i]);
    total += (uint64_t)npix[i] * sppc[i];
  }
  if (total >= (1ull << 40)) return fail(PT_ERR_INVALID, "pt_batch_render: too many path samples");
  if (!b->stream) {
    HIP_TRY(hipSetDevice(b->device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, b->device));
    b->n_cu = (uint32_t)std::max(1, prop.multiProcessorCount);
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  }
  // The scenes' uploads went through their own streams.
  for (size_t i = 0; i < ni; ++i) HIP_TRY(hipStreamSynchronize(b->insts[i]->scene->stream));

  uint32_t want = (bc && bc->wavefront) ? bc->wavefront : (b->tuning.wavefront ? b->tuning.wavefront : kDefaultWavefront);
  uint32_t cap = (uint32_t)std::min<uint64_t>(std::max(want, max_sppc), total);
  // Renders whose scenes use only the basic material kinds come first, so
  // that most chunks hold one class and shade with that class's kernel (the
  // basic set needs fewer registers); a chunk with any other kind shades
  // with every kind.
  std::vector<uint32_t> order;
  uint32_t kinds_basic = 0;
  for (int pass = 0; pass < 2; ++pass)
    for (uint32_t i = 0; i < (uint32_t)ni; ++i) {
      const uint32_t mk = b->insts[i]->scene->ds.mat_kinds;
      const bool basic = (mk & ~kMatBasic) == 0u;
      if (basic == (pass == 0)) order.push_back(i);
      if (basic) kinds_basic |= mk;
    }
  std::vector<ChunkSeg> segs;
  std::vector<ChunkInfo> chunks;
  plan_chunks(npix, sppc, order, cap, segs, chunks);
  std::vector<uint8_t> chunk_class(chunks.size(), 0);
  for (size_t c = 0; c < chunks.size(); ++c)
    for (uint32_t k = 0; k < chunks[c].n_segs; ++k)
      if (b->insts[segs[chunks[c].seg_begin + k].inst]->scene->ds.mat_kinds & ~kMatBasic) chunk_class[c] = 1;
  uint32_t max_paths = 0, max_pix = 0;
  for (const ChunkInfo &c : chunks) {
    max_paths = std::max(max_paths, c.paths);
    max_pix = std::max(max_pix, c.npix);
  }
  pt_status st = ensure_wave(b->ws, max_paths);
  if (st != PT_OK) return st;
  b->ws.wave.sort_keys = ((flags & PT_RENDER_SORT_HITS) || b->tuning.sort_bits) ? 1u : 0u;
  if (est == PT_EST_RESERVOIR) {
    st = ensure_reservoir(b->ws);
    if (st != PT_OK) return st;
  }

  // per-render tables
  std::vector<DevScene> scenes(ni);
  std::vector<RenderParams> params(ni);
  std::vector<DevCamera> cams(ni);
  std::vector<float> offs;
  std::vector<uint32_t> off_first(ni);
  std::vector<float *> films(ni);
  // Any-hit traversal of a batch reads one node format: the quantised nodes
  // only when every scene selects them.  On the seeded 64-render batch
  // (scenes of 1e4-5e5 triangles) the 128-byte nodes took 391.8 ms of shadow
  // rays per batch against 427.3 ms with quantised nodes for all (the
  // renders one by one, each scene with its own choice: 362.4 ms); a
  // per-lane choice of format spilled the any-hit kernel's registers and
  // took 792 ms.
  uint32_t batch_qany = 1u;
  for (size_t i = 0; i < ni; ++i) {
    BatchInst &in = *b->insts[i];
    scenes[i] = in.scene->ds;
    if (!in.scene->ds.qany) batch_qany = 0u;
    if (in.cfg.sampler == PT_SAMPLER_HALTON_PERM) {  // this render's own digit permutations
      if (in.perm_seed != in.cfg.seed || !in.perm.p) {
        st = build_perm(in.cfg.seed, in.perm, in.perm_off);
        if (st != PT_OK) return st;
        in.perm_seed = in.cfg.seed;
      }
      scenes[i].perm = in.perm.as<uint16_t>();
      scenes[i].perm_off = in.perm_off.as<uint32_t>();
      scenes[i].perm_seed = in.cfg.seed;
    }
    params[i] = make_params(&in.cfg, spp[i], sppc[i]);
    cams[i] = make_camera(in.scene->cam);
    off_first[i] = (uint32_t)offs.size();
    offs.resize(offs.size() + 2 * spp[i]);
    pixel_offsets(&in.cfg, spp[i], offs.data() + off_first[i]);
    const size_t fbytes = (size_t)npix[i] * 3 * sizeof(float);
    if (in.film_w != in.cfg.width || in.film_h != in.cfg.height || !(in.cfg.flags & PT_RENDER_ACCUMULATE) ||
        !in.film.p) {
      HIP_TRY(in.film.alloc(fbytes));
      HIP_TRY(hipMemsetAsync(in.film.p, 0, fbytes, b->stream));
      in.film_w = in.cfg.width;
      in.film_h = in.cfg.height;
      in.film_spp = 0;
    }
    films[i] = in.film.as<float>();
  }
  // Table uploads are synchronous: the previous render of this batch has
  // finished (pt_batch_render synchronises before returning).
  for (DevScene &sc : scenes) sc.qany = batch_qany;  // one shadow kernel variant for the batch
  HIP_TRY(upload_vec(b->d_scenes, scenes));
  HIP_TRY(upload_vec(b->d_params, params));
  HIP_TRY(upload_vec(b->d_cams, cams));
  HIP_TRY(upload_vec(b->d_offs, offs));
  HIP_TRY(upload_vec(b->d_off_first, off_first));
  HIP_TRY(upload_vec(b->d_films, films));
  HIP_TRY(upload_vec(b->d_segs, segs));
  HIP_TRY(upload_vec(b->d_chunks, chunks));
  HIP_TRY(b->d_cur.alloc(256));
  HIP_TRY(b->d_slots.alloc((size_t)3 * b->ws.cap * sizeof(uint32_t)));
  BatchView bv{};
  bv.scenes = b->d_scenes.as<DevScene>();
  bv.params = b->d_params.as<RenderParams>();
  bv.cams = b->d_cams.as<DevCamera>();
  bv.offs = b->d_offs.as<float>();
  bv.off_first = b->d_off_first.as<uint32_t>();
  bv.films = b->d_films.as<float *>();
  bv.segs = b->d_segs.as<ChunkSeg>();
  bv.chunks = b->d_chunks.as<ChunkInfo>();
  bv.cur = b->d_cur.as<uint32_t>();
  bv.slot_inst = b->d_slots.as<uint32_t>();
  bv.slot_pix = bv.slot_inst + b->ws.cap;
  bv.slot_k = bv.slot_pix + b->ws.cap;

  ChunkPlan plans[2]{};
  for (int k = 0; k < 2; ++k) {
    ChunkPlan &pl = plans[k];
    pl.max_paths = max_paths;
    pl.max_pix = max_pix;
    pl.nverts = nverts;
    pl.estimator = est;
    pl.mat_kinds = k == 0 ? kinds_basic : mat_kinds;
    pl.count = (flags & PT_RENDER_COUNTERS) != 0;
    pl.grids = make_grids(b->n_cu, b->tuning);
  }
  RenderParams rp0{};  // the kernels read the renders' own parameters
  rp0.estimator = est;
  rp0.chunk_paths = max_paths;
  const DevScene ds0 = scenes[0];
  const DevCamera cam0{};
  const bool use_graph = !(flags & PT_BATCH_NO_GRAPH) && !(flags & PT_RENDER_TIMING);

  HIP_TRY(hipMemsetAsync(b->ws.stats.p, 0, CNT_N * sizeof(unsigned long long), b->stream));
  bool used[2] = {false, false};
  for (uint8_t c : chunk_class) used[c] = true;
  for (int k = 0; k < 2 && use_graph; ++k) {
    if (!used[k]) continue;
    const ChunkPlan &pl = plans[k];
    // The graph's kernel arguments are fixed device addresses and sizes:
    // re-capture when any of them changed.
    std::vector<unsigned char> key(sizeof(pl) + sizeof(bv) + sizeof(Wave) + sizeof(DevScene));
    std::memcpy(key.data(), &pl, sizeof(pl));
    std::memcpy(key.data() + sizeof(pl), &bv, sizeof(bv));
    std::memcpy(key.data() + sizeof(pl) + sizeof(bv), &b->ws.wave, sizeof(Wave));
    std::memcpy(key.data() + sizeof(pl) + sizeof(bv) + sizeof(Wave), &ds0, sizeof(DevScene));
    if (key != b->graph_key[k] || !b->exec[k]) {
      if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
      if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
      b->exec[k] = nullptr;
      b->graph[k] = nullptr;
      b->graph_key[k].clear();
      HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      launch_chunk(b->stream, ds0, cam0, rp0, b->ws.wave, nullptr, nullptr, nullptr, nullptr, &bv, pl, nullptr);
      hipGraph_t g = nullptr;
      hipError_t e1 = hipGetLastError();
      hipError_t e2 = hipStreamEndCapture(b->stream, &g);
      if (e1 != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(PT_ERR_DEVICE, std::string("pt_batch_render: graph capture failed: ") +
                                       hipGetErrorString(e1 != hipSuccess ? e1 : e2));
      }
      b->graph[k] = g;
      HIP_TRY(hipGraphInstantiate(&b->exec[k], b->graph[k], nullptr, nullptr, 0));
      b->graph_key[k] = key;
    }
  }
  StageEvents ev;
  ev.on = (flags & PT_RENDER_TIMING) != 0;
  ev.st = b->stream;
  hipEvent_t tstart, tstop;
  HIP_TRY(hipEventCreate(&tstart));
  HIP_TRY(hipEventCreate(&tstop));
  HIP_TRY(hipEventRecord(tstart, b->stream));
  launch_set_chunk(b->stream, bv, 0u);
  for (size_t c = 0; c < chunks.size(); ++c) {
    const int k = chunk_class[c];
    if (use_graph)
      HIP_TRY(hipGraphLaunch(b->exec[k], b->stream));
    else
      launch_chunk(b->stream, ds0, cam0, rp0, b->ws.wave, nullptr, nullptr, nullptr, nullptr, &bv, plans[k], &ev);
  }
  HIP_TRY(hipEventRecord(tstop, b->stream));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < ni; ++i) b->insts[i]->film_spp += sppc[i];
  if (stats) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, tstart, tstop);
    unsigned long long c[CNT_N];
    HIP_TRY(hipMemcpy(c, b->ws.stats.p, sizeof(c), hipMemcpyDeviceToHost));
    fill_stats(stats, c, total, (uint64_t)chunks.size() * nverts, ms);
    ev.add_to(stats);
  }
  (void)hipEventDestroy(tstart);
  (void)hipEventDestroy(tstop);
  return PT_OK;
}

extern "C" pt_status pt_batch_render(pt_batch *b, const pt_batch_cfg *bc, pt_stats *stats) {
  if (!b) return fail(PT_ERR_INVALID, "pt_batch_render: NULL batch");
  return batch_render(b, bc, stats, 1u);
}

extern "C" pt_status pt_batch_film_read(pt_batch *b, uint32_t index, float *out) {
  if (!b || !out) return fail(PT_ERR_INVALID, "pt_batch_film_read: NULL argument");
  if (index >= b->insts.size()) return fail(PT_ERR_INVALID, "pt_batch_film_read: index out of range");
  BatchInst &in = *b->insts[index];
  if (!in.film.p || in.film_spp == 0) return fail(PT_ERR_INVALID, "pt_batch_film_read: nothing rendered");
  size_t n = (size_t)in.film_w * in.film_h * 3;
  HIP_TRY(hipMemcpy(out, in.film.p, n * sizeof(float), hipMemcpyDeviceToHost));
  float inv = 1.f / (float)in.film_spp;
  for (size_t i = 0; i < n; ++i) out[i] *= inv;
  return PT_OK;
}

// ------------------------------------------------------------------ autotuner
// Coordinate search: each tuned parameter in turn, the others held at their
// best so far; a candidate's time is the fastest of `repeats` renders of a
// 1/spp_div sample of the work.
static const uint32_t kTuneWavefront[] = {1u << 22, 1u << 23, 1u << 24, 1u << 25, 1u << 26};
static const uint32_t kTuneTrav[] = {4u, 6u, 8u};
static const uint32_t kTuneShade[] = {8u, 16u, 32u, 64u};
static const uint32_t kTuneLeaf[] = {1u, 2u, 4u, 8u};
static const uint32_t kTuneSort[] = {0u, 4u};

template <class F>
static pt_status tune_param(uint32_t *field, const uint32_t *cands, size_t nc, uint32_t repeats, F &&time_one,
                            double &best_ms) {
  uint32_t best = *field;
  for (size_t i = 0; i < nc; ++i) {
    *field = cands[i];
    double ms = 1e300;
    for (uint32_t r = 0; r < repeats; ++r) {
      double t = 0;
      pt_status st = time_one(t);
      if (st != PT_OK) return st;
      ms = std::min(ms, t);
    }
    if (ms < best_ms) {
      best_ms = ms;
      best = cands[i];
    }
  }
  *field = best;
  return PT_OK;
}

extern "C" pt_status pt_autotune(pt_scene *s, const pt_render_cfg *cfg, const pt_autotune_cfg *tc, pt_tuning *out) {
  uint32_t spp = 0, sppc = 0;
  pt_status st = check_cfg("pt_autotune", s, cfg, spp, sppc);
  if (st != PT_OK) return st;
  const uint32_t flags =
      (tc && tc->flags) ? tc->flags : (PT_TUNE_WAVEFRONT | PT_TUNE_GRIDS | PT_TUNE_LEAF | PT_TUNE_SORT);
  const uint32_t div = (tc && tc->spp_div) ? tc->spp_div : 8u;
  const uint32_t repeats = (tc && tc->repeats) ? tc->repeats : 1u;
  pt_render_cfg c = *cfg;
  c.spp_count = std::max(1u, sppc / div);
  c.wavefront = 0;
  c.flags &= ~(uint32_t)(PT_RENDER_ACCUMULATE | PT_RENDER_KEEP_SAMPLES | PT_RENDER_TIMING | PT_RENDER_COUNTERS |
                         PT_RENDER_SORT_HITS);
  pt_tuning &t = s->tuning;
  if (!t.wavefront) t.wavefront = kDefaultWavefront;
  if (!t.trav_blocks_per_cu) t.trav_blocks_per_cu = 8u;
  if (!t.shade_blocks_per_cu) t.shade_blocks_per_cu = 32u;
  auto time_one = [&](double &ms) -> pt_status {
    pt_stats sst;
    pt_status r = pt_render(s, &c, &sst);
    ms = sst.ms_total;
    return r;
  };
  double best = 1e300;
  {  // warm-up and the time of the starting point
    double ms = 0;
    st = time_one(ms);
    if (st != PT_OK) return st;
    st = time_one(ms);
    if (st != PT_OK) return st;
    best = ms;
  }
  if (flags & PT_TUNE_LEAF) {
    pt_build_cfg bcfg = s->build_cfg;
    uint32_t leaf = t.leaf_size ? t.leaf_size : 4u, best_leaf = leaf;
    for (uint32_t cand : kTuneLeaf) {
      if (cand == leaf) continue;
      bcfg.leaf_size = cand;
      st = pt_scene_commit(s, &bcfg);
      if (st != PT_OK) return st;
      double ms = 1e300;
      for (uint32_t r = 0; r < repeats; ++r) {
        double x = 0;
        st = time_one(x);
        if (st != PT_OK) return st;
        ms = std::min(ms, x);
      }
      if (ms < best) {
        best = ms;
        best_leaf = cand;
      }
    }
    bcfg.leaf_size = best_leaf;
    st = pt_scene_commit(s, &bcfg);
    if (st != PT_OK) return st;
  }
  if (flags & PT_TUNE_WAVEFRONT) {
    st = tune_param(&t.wavefront, kTuneWavefront, sizeof(kTuneWavefront) / 4, repeats, time_one, best);
    if (st != PT_OK) return st;
  }
  if (flags & PT_TUNE_GRIDS) {
    st = tune_param(&t.trav_blocks_per_cu, kTuneTrav, sizeof(kTuneTrav) / 4, repeats, time_one, best);
    if (st != PT_OK) return st;
    st = tune_param(&t.shade_blocks_per_cu, kTuneShade, sizeof(kTuneShade) / 4, repeats, time_one, best);
    if (st != PT_OK) return st;
  }
  if (flags & PT_TUNE_SORT) {
    st = tune_param(&t.sort_bits, kTuneSort, sizeof(kTuneSort) / 4, repeats, time_one, best);
    if (st != PT_OK) return st;
  }
  // the films of the tuning renders are not the caller's
  s->film_spp = 0;
  if (out) *out = t;
  return PT_OK;
}

extern "C" pt_status pt_batch_autotune(pt_batch *b, const pt_autotune_cfg *tc, pt_tuning *out) {
  if (!b) return fail(PT_ERR_INVALID, "pt_batch_autotune: NULL batch");
  const uint32_t flags =
      (tc && tc->flags) ? tc->flags : (PT_TUNE_WAVEFRONT | PT_TUNE_GRIDS | PT_TUNE_LEAF | PT_TUNE_SORT);
  const uint32_t div = (tc && tc->spp_div) ? tc->spp_div : 8u;
  const uint32_t repeats = (tc && tc->repeats) ? tc->repeats : 1u;
  pt_tuning &t = b->tuning;
  if (!t.wavefront) t.wavefront = kDefaultWavefront;
  if (!t.trav_blocks_per_cu) t.trav_blocks_per_cu = 8u;
  if (!t.shade_blocks_per_cu) t.shade_blocks_per_cu = 32u;
  pt_batch_cfg bc{};
  auto time_one = [&](double &ms) -> pt_status {
    pt_stats sst;
    pt_status r = batch_render(b, &bc, &sst, div);
    ms = sst.ms_total;
    return r;
  };
  double best = 1e300;
  {
    double ms = 0;
    pt_status st = time_one(ms);
    if (st != PT_OK) return st;
    st = time_one(ms);
    if (st != PT_OK) return st;
    best = ms;
  }
  pt_status st = PT_OK;
  if (flags & PT_TUNE_LEAF) {
    // one BVH leaf size for every scene of the batch: each candidate
    // re-commits the batch's distinct scenes (their other build options kept)
    std::vector<pt_scene *> uniq;
    for (auto &in : b->insts)
      if (std::find(uniq.begin(), uniq.end(), in->scene) == uniq.end()) uniq.push_back(in->scene);
    auto commit_all = [&](uint32_t leaf) -> pt_status {
      for (pt_scene *sc : uniq) {
        pt_build_cfg bcfg = sc->build_cfg;
        if ((bcfg.leaf_size ? bcfg.leaf_size : 4u) == leaf) continue;
        bcfg.leaf_size = leaf;
        pt_status r = pt_scene_commit(sc, &bcfg);
        if (r != PT_OK) return r;
      }
      return PT_OK;
    };
    const uint32_t leaf0 = t.leaf_size ? t.leaf_size : 4u;
    uint32_t best_leaf = leaf0;
    for (uint32_t cand : kTuneLeaf) {
      if (cand == leaf0) continue;
      st = commit_all(cand);
      if (st != PT_OK) return st;
      double ms = 1e300;
      for (uint32_t r = 0; r < repeats; ++r) {
        double x = 0;
        st = time_one(x);
        if (st != PT_OK) return st;
        ms = std::min(ms, x);
      }
      if (ms < best) {
        best = ms;
        best_leaf = cand;
      }
    }
    st = commit_all(best_leaf);
    if (st != PT_OK) return st;
    t.leaf_size = best_leaf;
  }
  if (flags & PT_TUNE_WAVEFRONT) {
    st = tune_param(&t.wavefront, kTuneWavefront, sizeof(kTuneWavefront) / 4, repeats, time_one, best);
    if (st != PT_OK) return st;
  }
  if (flags & PT_TUNE_GRIDS) {
    st = tune_param(&t.trav_blocks_per_cu, kTuneTrav, sizeof(kTuneTrav) / 4, repeats, time_one, best);
    if (st != PT_OK) return st;
    st = tune_param(&t.shade_blocks_per_cu, kTuneShade, sizeof(kTuneShade) / 4, repeats, time_one, best);
    if (st != PT_OK) return st;
  }
  if (flags & PT_TUNE_SORT) {
    st = tune_param(&t.sort_bits, kTuneSort, sizeof(kTuneSort) / 4, repeats, time_one, best);
    if (st != PT_OK) return st;
  }
  for (auto &in : b->insts) in->film_spp = 0;
  if (out) *out = t;
  return PT_OK;
}

// ------------------------------------------------------------------ parity entry points
namespace {
// Temporary device buffers of one parity call; each buffer lives in its own
// heap node so that growing the list never moves or frees a live allocation.
struct Scratch {
  std::vector<std::unique_ptr<DevBuf>> bufs;
  void *put(const void *src, size_t n) {
    bufs.push_back(std::make_unique<DevBuf>());
    DevBuf &b = *bufs.back();
    if (b.alloc(n) != hipSuccess) return nullptr;
    if (src && hipMemcpy(b.p, src, n, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return b.p;
  }
};
pt_scene *g_aux_scene() {
  static pt_scene *s = nullptr;
  if (!s) s = new pt_scene();
  return s;
}
}  // namespace

#define NEED(ptr)                                                                 \
  do {                                                                            \
    if (!(ptr)) return fail(PT_ERR_DEVICE, "device allocation or copy failed");   \
  } while (0)

extern "C" pt_status pt_sample(uint32_t sampler, uint32_t seed, uint32_t spp, const uint32_t *dim,
                               const uint64_t *index, const uint32_t *pixel, size_t n, float *out) {
  if (!dim || !index || !out) return fail(PT_ERR_INVALID, "pt_sample: NULL argument");
  if (sampler > PT_SAMPLER_STRATIFIED) return fail(PT_ERR_INVALID, "pt_sample: unknown sampler");
  if (n == 0) return PT_OK;
  for (size_t i = 0; i < n; ++i)
    if (index[i] >= (1ull << 32)) return fail(PT_ERR_INVALID, "pt_sample: index must be < 2^32");
  pt_scene *s = g_aux_scene();
  pt_status st = ensure_stream(s);
  if (st != PT_OK) return st;
  if (sampler == PT_SAMPLER_HALTON_PERM) {
    st = ensure_perm(s, seed);
    if (st != PT_OK) return st;
  }
  uint32_t sx = 1, sy = std::max(1u, spp);
  grid_factor(std::max(1u, spp), sx, sy);
  Scratch sc;
  auto *d_dim = (uint32_t *)sc.put(dim, n * 4);
  auto *d_idx = (uint64_t *)sc.put(index, n * 8);
  auto *d_pix = pixel ? (uint32_t *)sc.put(pixel, n * 4) : nullptr;
  auto *d_out = (float *)sc.put(nullptr, n * 4);
  NEED(d_dim); NEED(d_idx); NEED(d_out);
  if (pixel) NEED(d_pix);
  launch_sample(s->stream, s->ds, sampler, seed, sx, sy, d_dim, d_idx, d_pix, n, d_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(out, d_out, n * 4, hipMemcpyDeviceToHost));
  return PT_OK;
}

extern "C" pt_status pt_intersect(pt_scene *s, const float *rays, size_t n, int any_hit, int32_t *prim, float *t) {
  if (!s || !rays || !prim || !t) return fail(PT_ERR_INVALID, "pt_intersect: NULL argument");
  if (!s->committed) return fail(PT_ERR_INVALID, "pt_intersect: scene not committed");
  if (n == 0) return PT_OK;
  Scratch sc;
  auto *d_r = (float *)sc.put(rays, n * 32);
  auto *d_p = (int32_t *)sc.put(nullptr, n * 4);
  auto *d_t = (float *)sc.put(nullptr, n * 4);
  NEED(d_r); NEED(d_p); NEED(d_t);
  launch_intersect(s->stream, s->ds, d_r, n, any_hit, d_p, d_t);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(prim, d_p, n * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(t, d_t, n * 4, hipMemcpyDeviceToHost));
  return PT_OK;
}

extern "C" pt_status pt_bsdf_eval(const pt_material_desc *m, const float *wo, const float *wi, const float *nrm,
                                  const uint8_t *inside, size_t n, float *f, float *pdf) {
  if (!m || !wo || !wi || !nrm || !f || !pdf) return fail(PT_ERR_INVALID, "pt_bsdf_eval: NULL argument");
  if (!valid_kind(m->kind)) return fail(PT_ERR_INVALID, "pt_bsdf_eval: unknown material kind");
  if (n == 0) return PT_OK;
  pt_scene *s = g_aux_scene();
  pt_status st = ensure_stream(s);
  if (st != PT_OK) return st;
  Scratch sc;
  auto *a = (float *)sc.put(wo, n * 12);
  auto *b = (float *)sc.put(wi, n * 12);
  auto *c = (float *)sc.put(nrm, n * 12);
  auto *d = inside ? (uint8_t *)sc.put(inside, n) : nullptr;
  auto *of = (float *)sc.put(nullptr, n * 12);
  auto *op = (float *)sc.put(nullptr, n * 4);
  NEED(a); NEED(b); NEED(c); NEED(of); NEED(op);
  if (inside) NEED(d);
  launch_bsdf_eval(s->stream, *m, a, b, c, d, n, of, op);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(f, of, n * 12, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(pdf, op, n * 4, hipMemcpyDeviceToHost));
  return PT_OK;
}

extern "C" pt_status pt_bsdf_sample(const pt_material_desc *m, const float *wo, const float *nrm,
                                    const uint8_t *inside, const float *uc, const float *u2, const float *aux,
                                    size_t n, float *wi, float *f, float *pdf, uint32_t *flags, uint8_t *ok) {
  if (!m || !wo || !nrm || !uc || !u2 || !wi || !f || !pdf || !flags || !ok)
    return fail(PT_ERR_INVALID, "pt_bsdf_sample: NULL argument");
  if (!valid_kind(m->kind)) return fail(PT_ERR_INVALID, "pt_bsdf_sample: unknown material kind");
  if (n == 0) return PT_OK;
  pt_scene *s = g_aux_scene();
  pt_status st = ensure_stream(s);
  if (st != PT_OK) return st;
  Scratch sc;
  auto *a = (float *)sc.put(wo, n * 12);
  auto *c = (float *)sc.put(nrm, n * 12);
  auto *d = inside ? (uint8_t *)sc.put(inside, n) : nullptr;
  auto *e = (float *)sc.put(uc, n * 4);
  auto *g = (float *)sc.put(u2, n * 8);
  auto *x = aux ? (float *)sc.put(aux, n * 4) : nullptr;
  auto *owi = (float *)sc.put(nullptr, n * 12);
  auto *of = (float *)sc.put(nullptr, n * 12);
  auto *op = (float *)sc.put(nullptr, n * 4);
  auto *ofl = (uint32_t *)sc.put(nullptr, n * 4);
  auto *ook = (uint8_t *)sc.put(nullptr, n);
  NEED(a); NEED(c); NEED(e); NEED(g); NEED(owi); NEED(of); NEED(op); NEED(ofl); NEED(ook);
  if (inside) NEED(d);
  if (aux) NEED(x);
  launch_bsdf_sample(s->stream, *m, a, c, d, e, g, x, n, owi, of, op, ofl, ook);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(wi, owi, n * 12, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(f, of, n * 12, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(pdf, op, n * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(flags, ofl, n * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(ok, ook, n, hipMemcpyDeviceToHost));
  return PT_OK;
}

extern "C" pt_status pt_light_sample(pt_scene *s, uint32_t light, const float *u2, const float *ref, size_t n,
                                     float *p, float *nrm, float *pdf, float *le) {
  if (!s || !u2 || !ref || !p || !nrm || !pdf || !le) return fail(PT_ERR_INVALID, "pt_light_sample: NULL argument");
  if (!s->committed) return fail(PT_ERR_INVALID, "pt_light_sample: scene not committed");
  if (light >= s->lights.size()) return fail(PT_ERR_INVALID, "pt_light_sample: light index out of range");
  if (n == 0) return PT_OK;
  Scratch sc;
  auto *a = (float *)sc.put(u2, n * 8);
  auto *b = (float *)sc.put(ref, n * 12);
  auto *op = (float *)sc.put(nullptr, n * 12);
  auto *on = (float *)sc.put(nullptr, n * 12);
  auto *od = (float *)sc.put(nullptr, n * 4);
  auto *ol = (float *)sc.put(nullptr, n * 12);
  NEED(a); NEED(b); NEED(op); NEED(on); NEED(od); NEED(ol);
  launch_light_sample(s->stream, s->ds, light, a, b, n, op, on, od, ol);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(p, op, n * 12, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(nrm, on, n * 12, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(pdf, od, n * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(le, ol, n * 12, hipMemcpyDeviceToHost));
  return PT_OK;
}

extern "C" pt_status pt_env_sample(pt_scene *s, const float *u2, size_t n, float *dir, float *pdf) {
  if (!s || !u2 || !dir || !pdf) return fail(PT_ERR_INVALID, "pt_env_sample: NULL argument");
  if (!s->committed) return fail(PT_ERR_INVALID, "pt_env_sample: scene not committed");
  if (!s->ds.has_env) return fail(PT_ERR_INVALID, "pt_env_sample: the scene has no environment");
  if (n == 0) return PT_OK;
  Scratch sc;
  auto *a = (const float *)sc.put(u2, n * 8);
  auto *od = (float *)sc.put(nullptr, n * 12);
  auto *op = (float *)sc.put(nullptr, n * 4);
  NEED(a); NEED(od); NEED(op);
  launch_env_sample(s->stream, s->ds, a, n, od, op);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(dir, od, n * 12, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(pdf, op, n * 4, hipMemcpyDeviceToHost));
  return PT_OK;
}

extern "C" pt_status pt_device_synchronize(void) {
  if (!g_has_device()) return fail(PT_ERR_NO_DEVICE, "no HIP device visible");
  HIP_TRY(hipDeviceSynchronize());
  return PT_OK;
}
