/*
 * or_scene.c -- oracle restatement of the reference's geometry, BVH, lights,
 * environment texture and camera.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Pinning (tests/test_oracle_geometry.py): Moller-Trumbore KAT of
 * tests/test_mesh_io.cpp:26-32; the median-split mesh BVH against brute
 * force over every primitive; sphere/quad closed forms; quad and sphere
 * light sample/pdf closed forms; the equirect mapping round trip.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "or_scene.h"

#define GROW(arr, cnt, cap, T)                                                 \
  do {                                                                         \
    if ((cnt) >= (cap)) {                                                      \
      (cap) = (cap) ? 2 * (cap) : 16;                                          \
      (arr) = (T *)realloc((arr), sizeof(T) * (size_t)(cap));                  \
    }                                                                          \
  } while (0)

or_scene *or_scene_create(void) {
  or_scene *s = (or_scene *)calloc(1, sizeof(or_scene));
  return s;
}

void or_scene_destroy(or_scene *s) {
  if (!s) return;
  free(s->mats);
  free(s->tris);
  for (int i = 0; i < s->n_meshes; ++i) {
    free(s->meshes[i].nodes);
    free(s->meshes[i].leaf_ids);
    free(s->meshes[i].uv);
    free(s->meshes[i].uvidx);
  }
  for (int i = 0; i < s->n_tex; ++i) free(s->tex[i].rgb);
  free(s->tex);
  free(s->meshes);
  free(s->spheres);
  free(s->rects);
  free(s->curves);
  free(s->lights);
  free(s->env);
  free(s->env_func);
  free(s->env_cdf);
  free(s->env_marg);
  free(s->env_mcdf);
  free(s);
}

int or_scene_add_material(or_scene *s, const pt_material_desc *d) {
  GROW(s->mats, s->n_mats, s->cap_mats, pt_material_desc);
  s->mats[s->n_mats] = *d;
  return s->n_mats++;
}

int or_scene_add_mesh(or_scene *s, const float *pos, uint32_t nv, const int32_t *idx,
                      uint32_t nt, uint32_t mat, uint32_t flags) {
  return or_scene_add_mesh_normals(s, pos, nv, idx, nt, NULL, 0, NULL, mat, flags);
}

/* TriangleMesh(MeshData, material) (primitives/triangleMesh.cpp:5-26), with
 * optional MeshData::normals / normalIndices (mesh_io/meshData.hpp:6-13). */
int or_scene_add_mesh_normals(or_scene *s, const float *pos, uint32_t nv, const int32_t *idx,
                              uint32_t nt, const float *nrm, uint32_t nn, const int32_t *nidx,
                              uint32_t mat, uint32_t flags) {
  GROW(s->meshes, s->n_meshes, s->cap_meshes, or_mesh);
  or_mesh *m = &s->meshes[s->n_meshes];
  memset(m, 0, sizeof(*m));
  m->first_tri = s->n_tris;
  m->n_tris = (int)nt;
  m->mat = mat;
  m->flags = flags;
  m->seq = s->seq++;
  for (uint32_t t = 0; t < nt; ++t) {
    GROW(s->tris, s->n_tris, s->cap_tris, or_tri);
    or_tri *tr = &s->tris[s->n_tris++];
    for (int c = 0; c < 3; ++c) {
      int32_t vi = idx[3 * t + c];
      if (vi < 0 || (uint32_t)vi >= nv) return -1;
      tr->p[c] = V3(pos[3 * vi], pos[3 * vi + 1], pos[3 * vi + 2]);
    }
    tr->smooth = nidx != NULL;
    if (nidx) {
      for (int c = 0; c < 3; ++c) {
        int32_t ni = nidx[3 * t + c];
        if (ni < 0 || (uint32_t)ni >= nn) return -1;
        tr->n[c] = V3(nrm[3 * ni], nrm[3 * ni + 1], nrm[3 * ni + 2]);
      }
    }
    tr->mesh = s->n_meshes;
  }
  /* TriangleMesh::TriangleMesh: the box of all positions, -/+ gl::epsilon
   * (triangleMesh.cpp:10-18, constants.hpp:7) */
  v3 lo = vs(INFINITY), hi = vs(-INFINITY);
  for (uint32_t v = 0; v < nv; ++v) {
    v3 p = V3(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]);
    lo = V3(fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z));
    hi = V3(fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z));
  }
  m->box_lo = vsub(lo, vs(1e-3f));
  m->box_hi = vadd(hi, vs(1e-3f));
  return s->n_meshes++;
}

/* MeshData::uvs / uvIndices of mesh `mesh` (mesh_io/meshData.hpp:6-13) */
int or_scene_set_mesh_uv(or_scene *s, uint32_t mesh, const float *uv, uint32_t nuv, const int32_t *uvidx,
                         uint32_t nt) {
  if ((int)mesh >= s->n_meshes) return -1;
  or_mesh *m = &s->meshes[mesh];
  if ((int)nt != m->n_tris) return -1;
  for (size_t k = 0; k < 3 * (size_t)nt; ++k)
    if (uvidx[k] < 0 || (uint32_t)uvidx[k] >= nuv) return -1;
  free(m->uv);
  free(m->uvidx);
  m->uv = (float *)malloc(sizeof(float) * 2 * (nuv ? nuv : 1));
  m->uvidx = (int32_t *)malloc(sizeof(int32_t) * 3 * (nt ? nt : 1));
  memcpy(m->uv, uv, sizeof(float) * 2 * nuv);
  memcpy(m->uvidx, uvidx, sizeof(int32_t) * 3 * nt);
  return 0;
}

/* ImageTexture(file, ...) after loading (texture.cpp:3-95) */
int or_scene_add_texture_image(or_scene *s, const float *rgb, uint32_t w, uint32_t h) {
  if (w < 2 || h < 2) return -1;
  GROW(s->tex, s->n_tex, s->cap_tex, or_tex);
  or_tex *t = &s->tex[s->n_tex];
  memset(t, 0, sizeof(*t));
  t->kind = 0;
  t->w = (int)w;
  t->h = (int)h;
  t->rgb = (float *)malloc(sizeof(float) * 3 * (size_t)w * h);
  memcpy(t->rgb, rgb, sizeof(float) * 3 * (size_t)w * h);
  return s->n_tex++;
}

/* CheckerTexture(color1, color2, scale) (texture.cpp:151-156) */
int or_scene_add_texture_checker(or_scene *s, const float *c1, const float *c2, float scale) {
  GROW(s->tex, s->n_tex, s->cap_tex, or_tex);
  or_tex *t = &s->tex[s->n_tex];
  memset(t, 0, sizeof(*t));
  t->kind = 1;
  t->c1 = V3(c1[0], c1[1], c1[2]);
  t->c2 = V3(c2[0], c2[1], c2[2]);
  t->scale = scale;
  return s->n_tex++;
}

/* Texture2D::getTexelColor (texture.cpp:99-183).  ImageTexture, bilinear:
 * x = (uint)(u*w) (negative u converts to 0, as a saturating conversion
 * does), the fraction taken before x is clamped to w-2, likewise y.
 * CheckerTexture: color1 where sin(u scale) sin(v scale) < 0, else color2. */
v3 or_texture(const or_scene *s, uint32_t id, float u, float v) {
  const or_tex *t = &s->tex[id];
  if (t->kind == 1) {
    float sv = sinf(u * t->scale) * sinf(v * t->scale);
    return sv < 0.0f ? t->c1 : t->c2;
  }
  int W = t->w, H = t->h;
  float fx = u * (float)W, fy = v * (float)H;
  unsigned x = fx > 0.f ? (unsigned)fx : 0u, y = fy > 0.f ? (unsigned)fy : 0u;
  float x1 = fx - (float)x, y1 = fy - (float)y;
  if (x > (unsigned)(W - 2)) x = (unsigned)(W - 2);
  if (y > (unsigned)(H - 2)) y = (unsigned)(H - 2);
  float x2 = (float)(1.0 - x1), y2 = (float)(1.0 - y1);
#define TX(xx, yy) V3(t->rgb[3 * ((size_t)(yy) * W + (xx))], t->rgb[3 * ((size_t)(yy) * W + (xx)) + 1], t->rgb[3 * ((size_t)(yy) * W + (xx)) + 2])
  v3 c1 = TX(x, y), c2 = TX(x + 1, y), c3 = TX(x, y + 1), c4 = TX(x + 1, y + 1);
#undef TX
  v3 c12 = vadd(vscale(c1, x2), vscale(c2, x1));
  v3 c34 = vadd(vscale(c3, x2), vscale(c4, x1));
  return vadd(vscale(c12, y2), vscale(c34, y1));
}

int or_scene_add_sphere(or_scene *s, const float *c, float r, uint32_t mat, uint32_t flags) {
  GROW(s->spheres, s->n_spheres, s->cap_spheres, or_sphere);
  or_sphere *sp = &s->spheres[s->n_spheres];
  sp->c = V3(c[0], c[1], c[2]);
  sp->r = r;
  sp->mat = mat;
  sp->flags = flags;
  sp->seq = s->seq++;
  sp->light = -1;
  return s->n_spheres++;
}

int or_scene_add_rect(or_scene *s, int axis, float k, float d0min, float d0max, float d1min,
                      float d1max, uint32_t mat, uint32_t flags) {
  if (axis < 0 || axis > 2) return -1;
  GROW(s->rects, s->n_rects, s->cap_rects, or_rect);
  or_rect *r = &s->rects[s->n_rects];
  r->axis = axis;
  r->k = k;
  r->d0min = d0min; r->d0max = d0max; r->d1min = d1min; r->d1max = d1max;
  r->mat = mat;
  r->flags = flags;
  r->seq = s->seq++;
  r->light = -1;
  return s->n_rects++;
}

/* gl::evalBezier / evalBezierDerivative (primitives/bezier.hpp:12-35) */
static v3 bezier_at(const v3 *c, float u) {
  float u1 = 1.0f - u;
  float b0 = u1 * u1 * u1, b1 = 3.0f * u * u1 * u1, b2 = 3.0f * u * u * u1, b3 = u * u * u;
  return vadd(vadd(vadd(vscale(c[0], b0), vscale(c[1], b1)), vscale(c[2], b2)), vscale(c[3], b3));
}
static v3 bezier_deriv(const v3 *c, float u) {
  v3 d0 = vscale(vsub(c[1], c[0]), 3.0f), d1 = vscale(vsub(c[2], c[1]), 3.0f), d2 = vscale(vsub(c[3], c[2]), 3.0f);
  float u1 = 1.0f - u;
  return vadd(vadd(vscale(d0, u1 * u1), vscale(d1, 2.0f * u * u1)), vscale(d2, u * u));
}

int or_scene_add_curve(or_scene *s, const float *cps, float w0, float w1, uint32_t mat, uint32_t flags) {
  GROW(s->curves, s->n_curves, s->cap_curves, or_curve);
  or_curve *c = &s->curves[s->n_curves];
  for (int i = 0; i < 4; ++i) c->cps[i] = V3(cps[3 * i], cps[3 * i + 1], cps[3 * i + 2]);
  c->w0 = w0;
  c->w1 = w1;
  for (int i = 0; i <= 8; ++i) {  /* Curve::intersect (curve.hpp:60-66) */
    float u = (float)i / 8.0f;
    c->P[i] = bezier_at(c->cps, u);
    c->dP[i] = bezier_deriv(c->cps, u);
  }
  c->mat = mat;
  c->flags = flags & (PT_PRIM_FLIP_FACE | PT_PRIM_FORCE_INSIDE);
  return s->n_curves++;
}

int or_scene_set_env(or_scene *s, const float *rgb, uint32_t w, uint32_t h, float scale) {
  if (w < 2 || h < 2) return -1;
  free(s->env);
  s->env = (float *)malloc(sizeof(float) * 3 * (size_t)w * h);
  memcpy(s->env, rgb, sizeof(float) * 3 * (size_t)w * h);
  s->env_w = (int)w;
  s->env_h = (int)h;
  s->env_scale = scale;
  return 0;
}

int or_scene_set_medium(or_scene *s, const float *sigma_a, const float *sigma_s, float sigma_scale, float g) {
  if (!sigma_a) {
    s->has_medium = 0;
    return 0;
  }
  /* HomogeneousMedium::HomogeneousMedium (homogeneous.hpp:41-47) */
  s->med_sa = V3(sigma_a[0] * sigma_scale, sigma_a[1] * sigma_scale, sigma_a[2] * sigma_scale);
  s->med_ss = V3(sigma_s[0] * sigma_scale, sigma_s[1] * sigma_scale, sigma_s[2] * sigma_scale);
  s->med_g = g;
  s->has_medium = 1;
  return 0;
}

int or_scene_set_camera(or_scene *s, const pt_camera_desc *c) {
  s->cam = *c;
  s->has_cam = 1;
  /* PerspectiveCamera::PerspectiveCamera (camera.cpp:7-35) */
  float theta = c->fov / 2.f;
  float hh = tanf(theta);
  float ph = 1.0f * hh * 2.0f;
  float pw = ph * c->aspect;
  v3 up = V3(c->up[0], c->up[1], c->up[2]);
  v3 front = V3(c->front[0], c->front[1], c->front[2]);
  v3 lookat = vneg(vnorm(front));
  s->cam_h = vnorm(vcross(up, lookat));
  s->cam_v = vnorm(vcross(lookat, s->cam_h));
  s->cam_horiz = vscale(vscale(s->cam_h, pw), c->focal_dist);
  s->cam_vert = vscale(vscale(s->cam_v, ph), c->focal_dist);
  s->cam_bl_rel = vsub(vsub(vs(0.f), vdivs(s->cam_horiz, 2.0f)), vdivs(s->cam_vert, 2.0f));
  s->cam_bl_rel = vsub(s->cam_bl_rel, vscale(lookat, c->focal_dist));
  return 0;
}

/* ---------------- MeshBVHNode (primitives/meshBVH.cpp:6-55) ---------------- */
static const or_tri *g_sort_tris;
static int g_sort_axis;
static float centroid_axis(const or_tri *t, int axis) {
  v3 c = vdivs(vadd(vadd(t->p[0], t->p[1]), t->p[2]), 3.0f);
  return vget(c, axis);
}
static int cmp_centroid(const void *a, const void *b) {
  float ca = centroid_axis(&g_sort_tris[*(const int *)a], g_sort_axis);
  float cb = centroid_axis(&g_sort_tris[*(const int *)b], g_sort_axis);
  return (ca < cb) ? -1 : (cb < ca ? 1 : 0);
}
static int mesh_bvh_build(or_mesh *m, const or_tri *tris, int *ids, int start, int end) {
  GROW(m->nodes, m->n_nodes, m->cap_nodes, or_mnode);
  int ni = m->n_nodes++;
  v3 mn = vs(INFINITY), mx = vs(-INFINITY);
  for (int i = start; i < end; ++i) {
    const or_tri *t = &tris[ids[i]];
    for (int c = 0; c < 3; ++c) {
      mn = vmin3(mn, t->p[c]);
      mx = vmax3(mx, t->p[c]);
    }
  }
  m->nodes[ni].bmin = vsub(mn, vs(1e-3f)); /* gl::epsilon (constants.hpp:7) */
  m->nodes[ni].bmax = vadd(mx, vs(1e-3f));
  int count = end - start;
  if (count <= 4) {
    m->nodes[ni].left = -1;
    m->nodes[ni].right = -1;
    m->nodes[ni].leaf_first = m->n_leaf;
    m->nodes[ni].leaf_count = count;
    for (int i = start; i < end; ++i) {
      GROW(m->leaf_ids, m->n_leaf, m->cap_leaf, int);
      m->leaf_ids[m->n_leaf++] = ids[i];
    }
    return ni;
  }
  v3 ext = vsub(mx, mn);
  int axis = (ext.x > ext.y && ext.x > ext.z) ? 0 : (ext.y > ext.z ? 1 : 2);
  g_sort_tris = tris;
  g_sort_axis = axis;
  qsort(ids + start, (size_t)count, sizeof(int), cmp_centroid);
  int mid = count / 2;
  int l = mesh_bvh_build(m, tris, ids, start, start + mid);
  int r = mesh_bvh_build(m, tris, ids, start + mid, end);
  m->nodes[ni].left = l;
  m->nodes[ni].right = r;
  m->nodes[ni].leaf_count = 0;
  return ni;
}

/* ---------------- environment importance-sampling tables (new; the reference
 * samples the environment uniformly, texture.hpp:120-131) ---------------- */
static float texel_max(const or_scene *s, int x, int y) {
  const float *c = &s->env[3 * ((size_t)y * s->env_w + x)];
  return fmaxf(c[0], fmaxf(c[1], c[2]));
}
static void env_build_cdf(or_scene *s) {
  int W = s->env_w, H = s->env_h;
  s->env_func = (float *)malloc(sizeof(float) * (size_t)W * H);
  s->env_cdf = (float *)malloc(sizeof(float) * (size_t)(W + 1) * H);
  s->env_marg = (float *)malloc(sizeof(float) * (size_t)H);
  s->env_mcdf = (float *)malloc(sizeof(float) * (size_t)(H + 1));
  for (int y = 0; y < H; ++y) {
    int y0 = y < H - 2 ? y : H - 2;
    float st = sinf((float)M_PI * ((float)y + 0.5f) / (float)H);
    for (int x = 0; x < W; ++x) {
      int x0 = x < W - 2 ? x : W - 2;
      float v = fmaxf(fmaxf(texel_max(s, x0, y0), texel_max(s, x0 + 1, y0)),
                      fmaxf(texel_max(s, x0, y0 + 1), texel_max(s, x0 + 1, y0 + 1)));
      s->env_func[(size_t)y * W + x] = v * st;
    }
  }
  for (int y = 0; y < H; ++y) {
    float *cdf = &s->env_cdf[(size_t)y * (W + 1)];
    const float *f = &s->env_func[(size_t)y * W];
    double acc = 0.0;
    cdf[0] = 0.f;
    for (int x = 0; x < W; ++x) {
      acc += (double)f[x] / W;
      cdf[x + 1] = (float)acc;
    }
    s->env_marg[y] = (float)acc;
    if (acc > 0.0)
      for (int x = 1; x <= W; ++x) cdf[x] = (float)((double)cdf[x] / acc);
    else
      for (int x = 1; x <= W; ++x) cdf[x] = (float)x / (float)W;
  }
  double acc = 0.0;
  s->env_mcdf[0] = 0.f;
  for (int y = 0; y < H; ++y) {
    acc += (double)s->env_marg[y] / H;
    s->env_mcdf[y + 1] = (float)acc;
  }
  s->env_integral = (float)acc;
  for (int y = 1; y <= H; ++y) s->env_mcdf[y] = acc > 0.0 ? (float)((double)s->env_mcdf[y] / acc) : (float)y / H;
}

int or_scene_commit(or_scene *s) {
  for (int mi = 0; mi < s->n_meshes; ++mi) {
    or_mesh *m = &s->meshes[mi];
    if (m->n_tris == 0) continue;
    int *ids = (int *)malloc(sizeof(int) * (size_t)m->n_tris);
    for (int i = 0; i < m->n_tris; ++i) ids[i] = m->first_tri + i;
    mesh_bvh_build(m, s->tris, ids, 0, m->n_tris);
    free(ids);
  }
  /* discover_emissive_objects_as_lights (base/lightDiscovery.hpp:5-81): rects
   * and spheres with an emissive material of positive intensity, in object
   * insertion order; the environment light is appended last
   * (renderManager.hpp:92-95). */
  free(s->lights);
  s->lights = (or_light *)calloc((size_t)(s->n_rects + s->n_spheres + 1), sizeof(or_light));
  s->n_lights = 0;
  for (uint32_t q = 0; q < s->seq; ++q) {
    for (int i = 0; i < s->n_rects; ++i) {
      or_rect *r = &s->rects[i];
      if (r->seq != q) continue;
      const pt_material_desc *md = &s->mats[r->mat];
      if (md->kind != PT_MAT_EMITTER || !(md->p[3] > 0.f) || (r->flags & PT_PRIM_NOT_LIGHT)) continue;
      or_light *L = &s->lights[s->n_lights];
      L->type = OR_LIGHT_QUAD;
      /* QuadLight(AARectangle<Axis>) vertex order (light.hpp:71-108) */
      v3 v[4];
      float k = r->k;
      if (r->axis == 0) {
        v[0] = V3(k, r->d0min, r->d1min); v[1] = V3(k, r->d0max, r->d1min);
        v[2] = V3(k, r->d0max, r->d1max); v[3] = V3(k, r->d0min, r->d1max);
      } else if (r->axis == 1) {
        v[0] = V3(r->d0min, k, r->d1min); v[1] = V3(r->d0max, k, r->d1min);
        v[2] = V3(r->d0max, k, r->d1max); v[3] = V3(r->d0min, k, r->d1max);
      } else {
        v[0] = V3(r->d0min, r->d1min, k); v[1] = V3(r->d0max, r->d1min, k);
        v[2] = V3(r->d0max, r->d1max, k); v[3] = V3(r->d0min, r->d1max, k);
      }
      for (int c = 0; c < 4; ++c) L->v[c] = v[c];
      v3 e1 = vsub(v[1], v[0]), e2 = vsub(v[3], v[0]);
      L->n = vnorm(vcross(e1, e2));
      L->area = vlen(vcross(e1, e2));
      L->le = vscale(V3(md->p[0], md->p[1], md->p[2]), md->p[3]);
      L->src = i;
      r->light = s->n_lights++;
    }
    for (int i = 0; i < s->n_spheres; ++i) {
      or_sphere *sp = &s->spheres[i];
      if (sp->seq != q) continue;
      const pt_material_desc *md = &s->mats[sp->mat];
      if (md->kind != PT_MAT_EMITTER || !(md->p[3] > 0.f) || (sp->flags & PT_PRIM_NOT_LIGHT)) continue;
      or_light *L = &s->lights[s->n_lights];
      L->type = OR_LIGHT_SPHERE;
      L->c = sp->c;
      L->r = sp->r;
      L->area = (float)(4 * M_PI * sp->r * sp->r);
      L->le = vscale(V3(md->p[0], md->p[1], md->p[2]), md->p[3]);
      L->src = i;
      sp->light = s->n_lights++;
    }
  }
  if (s->env) {
    or_light *L = &s->lights[s->n_lights++];
    L->type = OR_LIGHT_ENV;
    free(s->env_func); free(s->env_cdf); free(s->env_marg); free(s->env_mcdf);
    env_build_cdf(s);
  }
  s->committed = 1;
  return 0;
}

int or_scene_num_lights(const or_scene *s) { return s->n_lights; }

/* ---------------- intersection ---------------- */

/* HitRecord::set_normal (material/materialBase.hpp:48-52) plus the
 * wrapper semantics: Translate/Rotate/Scale force is_inside true
 * (objectTransform.hpp:60,156; objectTransform.cpp:132,172,211), then a
 * FlipFace outside them toggles it (objectTransform.hpp:21). */
static void set_normal(or_hit *h, v3 d, v3 n, uint32_t flags) {
  h->inside = vdot(d, n) < 0.f;
  h->n = h->inside ? n : vneg(n);
  if (flags & PT_PRIM_FORCE_INSIDE) h->inside = 1;
  if (flags & PT_PRIM_FLIP_FACE) h->inside = !h->inside;
}

/* AABB::intersect (utils/aabb.hpp:21-42): Kensler slab on the raw direction */
static int aabb_hit(v3 bmin, v3 bmax, v3 o, v3 d, float tmin, float tmax) {
  for (int i = 0; i < 3; ++i) {
    float inv = 1.0f / vget(d, i);
    float t0 = (vget(bmin, i) - vget(o, i)) * inv;
    float t1 = (vget(bmax, i) - vget(o, i)) * inv;
    if (inv < 0.0f) { float tt = t0; t0 = t1; t1 = tt; }
    tmin = t0 > tmin ? t0 : tmin;
    tmax = t1 < tmax ? t1 : tmax;
    if (tmax <= tmin) return 0;
  }
  return 1;
}

/* MeshBVHNode::hitTriangle (meshBVH.cpp:85-143), geometric normal. */
static int hit_triangle(const or_tri *tr, v3 o, v3 d, float tmin, float tmax, float *tout) {
  v3 e1 = vsub(tr->p[1], tr->p[0]);
  v3 e2 = vsub(tr->p[2], tr->p[0]);
  v3 P = vcross(d, e2);
  float det = vdot(e1, P);
  if (fabsf(det) < 1e-8f) return 0;
  float inv_det = 1.0f / det;
  v3 T = vsub(o, tr->p[0]);
  float u = vdot(T, P) * inv_det;
  if (u < 0.f || u > 1.f) return 0;
  v3 Q = vcross(T, e1);
  float v = vdot(d, Q) * inv_det;
  if (v < 0.f || u + v > 1.f) return 0;
  float t = vdot(e2, Q) * inv_det;
  if (t < tmin || t > tmax) return 0;
  *tout = t;
  return 1;
}

/* MeshBVHNode::intersect (meshBVH.cpp:57-83): left, then right with the
 * shrunk bound; leaf triangles in order, inclusive bounds (later wins ties). */
static int mesh_hit(const or_scene *s, const or_mesh *m, int ni, v3 o, v3 d, float tmin,
                    float tmax, float *t, int *tri, int any, uint64_t *cnt) {
  const or_mnode *nd = &m->nodes[ni];
  if (cnt) cnt[0]++;
  if (!aabb_hit(nd->bmin, nd->bmax, o, d, tmin, tmax)) return 0;
  int hit = 0;
  float closest = tmax;
  if (nd->left < 0) {
    for (int i = 0; i < nd->leaf_count; ++i) {
      int id = m->leaf_ids[nd->leaf_first + i];
      float tt;
      if (cnt) cnt[1]++;
      if (hit_triangle(&s->tris[id], o, d, tmin, closest, &tt)) {
        hit = 1;
        closest = tt;
        *t = tt;
        *tri = id;
        if (any) return 1;
      }
    }
  } else {
    if (mesh_hit(s, m, nd->left, o, d, tmin, closest, t, tri, any, cnt)) {
      hit = 1;
      closest = *t;
      if (any) return 1;
    }
    if (mesh_hit(s, m, nd->right, o, d, tmin, closest, t, tri, any, cnt)) hit = 1;
  }
  return hit;
}

/* Sphere::intersect (base/primitive.hpp:124-170): normalised direction */
static int sphere_hit(const or_sphere *sp, v3 o, v3 d, float tmin, float tmax, float *tout) {
  v3 dn = vnorm(d);
  v3 oc = vsub(o, sp->c);
  float a = vdot(dn, dn);
  float b = 2.f * vdot(dn, oc);
  float c = vdot(oc, oc) - sp->r * sp->r;
  float disc = b * b - 4.f * a * c;
  if (disc < 0.f) return 0;
  float delta = sqrtf(disc);
  float t = (-b - delta) / (2.f * a);
  if (t < tmin || t > tmax) {
    t = (-b + delta) / (2.f * a);
    if (t < tmin || t > tmax) return 0;
  }
  *tout = t;
  return 1;
}

/* AARectangle<Axis>::intersect (base/primitive.cpp:31-122) */
static int rect_hit(const or_rect *r, v3 o, v3 d, float tmin, float tmax, float *tout) {
  v3 dn = vnorm(d);
  int a = r->axis, b0 = (a == 0) ? 1 : 0, b1 = (a == 2) ? 1 : 2;
  float t = (r->k - vget(o, a)) / vget(dn, a);
  if (t < tmin || t > tmax) return 0;
  float d0 = vget(o, b0) + t * vget(dn, b0);
  float d1 = vget(o, b1) + t * vget(dn, b1);
  if (d0 < r->d0min || d0 > r->d0max || d1 < r->d1min || d1 > r->d1max) return 0;
  *tout = t;
  return 1;
}

/* Closest hit over the scene: every object in turn with a shrinking bound
 * (the reference's BVHNode::intersect, bvh.hpp:27-41, visits objects in a
 * different order; closest-hit results agree except for exact ties). */
/* Curve::intersect (primitives/curve.hpp:47-130): 8 cylinder segments, the
 * closest root whose projection lies on its segment; solveQuadratic
 * (utils/utility.hpp:112-124).  Returns the segment and its hit. */
static int curve_hit(const or_curve *c, v3 o, v3 d, float tmin, float tmax, float *tout, int *seg_out,
                     v3 *n_out, v3 *tan_out, v3 *p_out) {
  float best = tmax;
  int found = 0;
  for (int i = 0; i < 8; ++i) {
    v3 pa = c->P[i], pb = c->P[i + 1];
    float u0 = (float)i / 8.0f, u1 = (float)(i + 1) / 8.0f;
    float r0 = c->w0 + (c->w1 - c->w0) * u0, r1 = c->w0 + (c->w1 - c->w0) * u1;
    float radius = 0.5f * (r0 + r1);
    v3 axis = vnorm(vsub(pb, pa));
    float len = vlen(vsub(pb, pa));
    v3 dp = vsub(d, vscale(axis, vdot(d, axis)));
    v3 de = vsub(o, pa);
    v3 op = vsub(de, vscale(axis, vdot(de, axis)));
    float A = vdot(dp, dp), B = 2.0f * vdot(dp, op), C = vdot(op, op) - radius * radius;
    float disc = B * B - 4.f * A * C;
    if (disc < 0.f) continue;
    float sq = sqrtf(disc);
    float q = B < 0.f ? -0.5f * (B - sq) : -0.5f * (B + sq);
    float roots[2] = {q / A, C / q};
    if (roots[0] > roots[1]) { float x = roots[0]; roots[0] = roots[1]; roots[1] = x; }
    int nr = roots[0] == roots[1] ? 1 : 2;
    for (int r = 0; r < nr; ++r) {
      float t = roots[r];
      if (t < tmin || t > best) continue;
      v3 hp = vadd(o, vscale(d, t));
      float proj = vdot(vsub(hp, pa), axis);
      if (proj < 0.f || proj > len) continue;
      best = t;
      found = 1;
      *seg_out = i;
      *p_out = hp;
      float uu = proj / len;
      v3 tg = vnorm(vadd(c->dP[i], vscale(vsub(c->dP[i + 1], c->dP[i]), uu)));
      v3 v = vsub(hp, vadd(pa, vscale(axis, proj)));
      *tan_out = tg;
      *n_out = vnorm(vsub(v, vscale(tg, vdot(v, tg))));
    }
  }
  if (found) *tout = best;
  return found;
}

/* A textured material (pt_material_desc.flags bit 0) takes its colour from
 * texture flags >> 8 at the hit's texture coordinates. */
static void hit_texture(const or_scene *s, or_hit *h) {
  const uint32_t fl = s->mats[h->mat].flags;
  h->has_tex = (fl & 1u) != 0;
  h->tex = h->has_tex ? or_texture(s, fl >> 8, h->u, h->v) : vs(0.f);
}

int or_scene_intersect(const or_scene *s, v3 o, v3 d, float tmin, float tmax, int any,
                       or_hit *h, uint64_t *cnt) {
  float best = tmax;
  int kind = -1, idx = -1;
  for (int mi = 0; mi < s->n_meshes; ++mi) {
    const or_mesh *m = &s->meshes[mi];
    if (m->n_tris == 0) continue;
    float t;
    int tri;
    if (mesh_hit(s, m, 0, o, d, tmin, best, &t, &tri, any, cnt)) {
      best = t; kind = 0; idx = tri;
      if (any) return 1;
    }
  }
  for (int i = 0; i < s->n_spheres; ++i) {
    float t;
    if (cnt) cnt[1]++;
    if (sphere_hit(&s->spheres[i], o, d, tmin, best, &t)) {
      best = t; kind = 1; idx = i;
      if (any) return 1;
    }
  }
  for (int i = 0; i < s->n_rects; ++i) {
    float t;
    if (cnt) cnt[1]++;
    if (rect_hit(&s->rects[i], o, d, tmin, best, &t)) {
      best = t; kind = 2; idx = i;
      if (any) return 1;
    }
  }
  int cseg = 0;
  v3 cn = vs(0.f), ctan = vs(0.f), cp = vs(0.f);
  for (int i = 0; i < s->n_curves; ++i) {
    float t;
    int sg;
    v3 nn, tt, pp;
    if (cnt) cnt[1] += 8;
    if (curve_hit(&s->curves[i], o, d, tmin, best, &t, &sg, &nn, &tt, &pp)) {
      best = t; kind = 3; idx = i; cseg = sg; cn = nn; ctan = tt; cp = pp;
      if (any) return 1;
    }
  }
  if (kind < 0) return 0;
  if (!h) return 1;
  h->t = best;
  h->ng_light = vs(0.f);
  h->tangent = vs(0.f);
  h->u = h->v = 0.f;  /* Curve sets no texCoords */
  if (kind == 3) {
    const or_curve *c = &s->curves[idx];
    h->p = cp;
    set_normal(h, d, cn, c->flags);
    h->tangent = ctan;
    h->mat = c->mat;
    h->prim = s->n_tris + s->n_spheres + s->n_rects + 8 * idx + cseg;
    hit_texture(s, h);
    return 1;
  }
  if (kind == 0) {
    const or_tri *tr = &s->tris[idx];
    v3 e1 = vsub(tr->p[1], tr->p[0]), e2 = vsub(tr->p[2], tr->p[0]);
    v3 ng = vnorm(vcross(e1, e2));
    h->p = vadd(o, vscale(d, best));
    const or_mesh *m = &s->meshes[tr->mesh];
    v3 ns = ng;
    if (tr->smooth) {
      /* MeshBVHNode::hitTriangle (meshBVH.cpp:95-130): barycentrics of the
       * hit, interpolated normal, turned to the geometric normal's side */
      v3 P = vcross(d, e2);
      float inv = 1.0f / vdot(e1, P);
      v3 T = vsub(o, tr->p[0]);
      float u = vdot(T, P) * inv;
      v3 Q = vcross(T, e1);
      float v = vdot(d, Q) * inv;
      ns = vnorm(vadd(vadd(vscale(tr->n[0], 1.f - u - v), vscale(tr->n[1], u)), vscale(tr->n[2], v)));
      if (vdot(ns, ng) < 0.f) ns = vneg(ns);
    }
    set_normal(h, d, ns, m->flags);
    h->mat = m->mat;
    h->prim = idx;
    if (m->uv) {
      /* MeshBVHNode::hitTriangle (meshBVH.cpp:133-138): the hit's
       * barycentrics, then (1-u-v) uv0 + u uv1 + v uv2 */
      v3 P = vcross(d, e2);
      float inv = 1.0f / vdot(e1, P);
      v3 T = vsub(o, tr->p[0]);
      float bu = vdot(T, P) * inv;
      v3 Q = vcross(T, e1);
      float bv = vdot(d, Q) * inv;
      const int32_t *ui = &m->uvidx[3 * (size_t)(idx - m->first_tri)];
      const float *u0 = &m->uv[2 * (size_t)ui[0]], *u1 = &m->uv[2 * (size_t)ui[1]], *u2 = &m->uv[2 * (size_t)ui[2]];
      float w0 = 1.f - bu - bv;
      h->u = w0 * u0[0] + bu * u1[0] + bv * u2[0];
      h->v = w0 * u0[1] + bu * u1[1] + bv * u2[1];
    } else {
      /* TriangleMesh's planar fallback (triangleMesh.cpp:38-51): the axes
       * the normal points least along, scaled to the mesh box, clamped */
      v3 n = V3(fabsf(h->n.x), fabsf(h->n.y), fabsf(h->n.z));
      int au = (n.x > n.y && n.x > n.z) ? 1 : 0;
      int av = (au == 1) ? 2 : (n.y > n.z ? 2 : 1);
      float uu = (vget(h->p, au) - vget(m->box_lo, au)) / (vget(m->box_hi, au) - vget(m->box_lo, au));
      float vv = (vget(h->p, av) - vget(m->box_lo, av)) / (vget(m->box_hi, av) - vget(m->box_lo, av));
      h->u = fminf(fmaxf(uu, 0.f), 1.f);
      h->v = fminf(fmaxf(vv, 0.f), 1.f);
    }
  } else if (kind == 1) {
    const or_sphere *sp = &s->spheres[idx];
    v3 dn = vnorm(d);
    h->p = vadd(o, vscale(dn, best));
    v3 nn = vnorm(vsub(h->p, sp->c));
    set_normal(h, d, nn, sp->flags);
    h->ng_light = nn;
    h->mat = sp->mat;
    h->prim = s->n_tris + idx;
    /* Sphere::intersect's texCoords (primitive.hpp:157-163) */
    v3 q = vnorm(vsub(h->p, sp->c));
    float phi = atan2f(q.z, q.x), theta = asinf(q.y);
    float u = (float)(1 - (phi + M_PI) / (2 * M_PI));
    h->v = (float)((theta + M_PI / 2) / M_PI);
    h->u = fmodf(u + 0.75f, 1.0f);
  } else {
    const or_rect *r = &s->rects[idx];
    v3 dn = vnorm(d);
    h->p = vadd(o, vscale(dn, best));
    v3 nn = r->axis == 0 ? V3(1, 0, 0) : (r->axis == 1 ? V3(0, 1, 0) : V3(0, 0, 1));
    set_normal(h, d, nn, r->flags);
    h->ng_light = (r->light >= 0) ? s->lights[r->light].n : nn;
    h->mat = r->mat;
    h->prim = s->n_tris + s->n_spheres + idx;
    /* AARectangle's texCoords (primitive.cpp:55-57, 86-88, 117-119) */
    float d0 = r->axis == 0 ? h->p.y : h->p.x, d1 = r->axis == 2 ? h->p.y : h->p.z;
    h->u = (d0 - r->d0min) / (r->d0max - r->d0min);
    h->v = (d1 - r->d1min) / (r->d1max - r->d1min);
  }
  hit_texture(s, h);
  return 1;
}

/* Texture `id` at n (u, v) pairs: out n*3 (tests). */
int or_texture_batch(const or_scene *s, uint32_t id, const float *uv, size_t n, float *out) {
  if ((int)id >= s->n_tex) return -1;
  for (size_t i = 0; i < n; ++i) {
    v3 c = or_texture(s, id, uv[2 * i], uv[2 * i + 1]);
    out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
  }
  return 0;
}

/* HitRecord::texCoords of the closest hits of n rays (n*8 floats): uv n*2,
 * (-1, -1) on a miss (tests). */
int or_hit_uv_batch(const or_scene *s, const float *rays, size_t n, float *uv) {
  for (size_t i = 0; i < n; ++i) {
    const float *r = &rays[8 * i];
    or_hit h;
    memset(&h, 0, sizeof(h));
    if (or_scene_intersect(s, V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]), r[6], r[7], 0, &h, NULL)) {
      uv[2 * i] = h.u;
      uv[2 * i + 1] = h.v;
    } else {
      uv[2 * i] = uv[2 * i + 1] = -1.f;
    }
  }
  return 0;
}

int or_prim_light(const or_scene *s, int prim) {
  if (prim < s->n_tris) return -1;
  prim -= s->n_tris;
  if (prim < s->n_spheres) return s->spheres[prim].light;
  prim -= s->n_spheres;
  if (prim >= s->n_rects) return -1;  /* curves are never lights */
  return s->rects[prim].light;
}

int or_intersect_batch(const or_scene *s, const float *rays, size_t n, int any_hit,
                       int32_t *prim, float *t) {
#pragma omp parallel for schedule(dynamic, 256)
  for (long long i = 0; i < (long long)n; ++i) {
    const float *r = &rays[8 * i];
    or_hit h;
    int hit = or_scene_intersect(s, V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]), r[6], r[7],
                                 any_hit, any_hit ? NULL : &h, NULL);
    if (any_hit) {
      prim[i] = hit ? 1 : -1;
      t[i] = 0.f;
    } else {
      prim[i] = hit ? h.prim : -1;
      t[i] = hit ? h.t : INFINITY;
    }
  }
  return 0;
}

/* Closest hit's shading normal (turned against the ray, HitRecord::set_normal)
 * and is_inside, for rays that hit; zeros and -1 for misses. */
int or_surface_batch(const or_scene *s, const float *rays, size_t n, float *nrm, int32_t *inside) {
#pragma omp parallel for schedule(dynamic, 256)
  for (long long i = 0; i < (long long)n; ++i) {
    const float *r = &rays[8 * i];
    or_hit h;
    int hit = or_scene_intersect(s, V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]), r[6], r[7], 0, &h, NULL);
    nrm[3 * i] = hit ? h.n.x : 0.f;
    nrm[3 * i + 1] = hit ? h.n.y : 0.f;
    nrm[3 * i + 2] = hit ? h.n.z : 0.f;
    inside[i] = hit ? h.inside : -1;
  }
  return 0;
}

/* Brute force over every primitive (no BVH), closest hit with later-wins ties. */
int or_brute_intersect_batch(const or_scene *s, const float *rays, size_t n, int32_t *prim,
                             float *t) {
#pragma omp parallel for schedule(dynamic, 64)
  for (long long i = 0; i < (long long)n; ++i) {
    const float *r = &rays[8 * i];
    v3 o = V3(r[0], r[1], r[2]), d = V3(r[3], r[4], r[5]);
    float best = r[7];
    int id = -1;
    for (int k = 0; k < s->n_tris; ++k) {
      float tt;
      if (hit_triangle(&s->tris[k], o, d, r[6], best, &tt)) { best = tt; id = k; }
    }
    for (int k = 0; k < s->n_spheres; ++k) {
      float tt;
      if (sphere_hit(&s->spheres[k], o, d, r[6], best, &tt)) { best = tt; id = s->n_tris + k; }
    }
    for (int k = 0; k < s->n_rects; ++k) {
      float tt;
      if (rect_hit(&s->rects[k], o, d, r[6], best, &tt)) {
        best = tt;
        id = s->n_tris + s->n_spheres + k;
      }
    }
    for (int k = 0; k < s->n_curves; ++k) {
      float tt;
      int sg;
      v3 a, b, c;
      if (curve_hit(&s->curves[k], o, d, r[6], best, &tt, &sg, &a, &b, &c)) {
        best = tt;
        id = s->n_tris + s->n_spheres + s->n_rects + 8 * k + sg;
      }
    }
    prim[i] = id;
    t[i] = id >= 0 ? best : INFINITY;
  }
  return 0;
}

/* ---------------- lights ---------------- */

/* HDRITexture::getTexelColor(direction) (texture.hpp:105-115) and
 * ImageTexture::getTexelColor BILINEAR (texture.cpp:111-142): no wrap, the
 * texel index is clamped to size-2 after the fractions are taken. */
v3 or_env_lookup(const or_scene *s, v3 dir) {
  v3 dn = vnorm(dir);
  float theta = acosf(-dn.y);
  float phi = (float)(atan2f(-dn.z, dn.x) + M_PI);
  float u = (float)(phi / (2.0f * M_PI));
  float v = (float)(theta / M_PI);
  int W = s->env_w, H = s->env_h;
  float fx = u * (float)W, fy = v * (float)H;
  unsigned x = (unsigned)fx, y = (unsigned)fy;
  float x1 = fx - (float)x, y1 = fy - (float)y;
  if (x > (unsigned)(W - 2)) x = (unsigned)(W - 2);
  if (y > (unsigned)(H - 2)) y = (unsigned)(H - 2);
  float x2 = 1.0f - x1, y2 = 1.0f - y1;
#define TX(xx, yy) V3(s->env[3 * ((size_t)(yy) * W + (xx))], s->env[3 * ((size_t)(yy) * W + (xx)) + 1], s->env[3 * ((size_t)(yy) * W + (xx)) + 2])
  v3 c1 = TX(x, y), c2 = TX(x + 1, y), c3 = TX(x, y + 1), c4 = TX(x + 1, y + 1);
#undef TX
  v3 c12 = vadd(vscale(c1, x2), vscale(c2, x1));
  v3 c34 = vadd(vscale(c3, x2), vscale(c4, x1));
  v3 c = vadd(vscale(c12, y2), vscale(c34, y1));
  return vscale(c, s->env_scale);
}

/* ImageTexture::getTexelColor at texture coordinates (u, v) (texture.cpp:
 * 111-142, bilinear with the clamp to w-2/h-2), times the environment scale:
 * the lookup EnvironmentLight::texture->getTexelColor(offset) makes. */
v3 or_env_lookup_uv(const or_scene *s, float u, float v) {
  int W = s->env_w, H = s->env_h;
  float fx = u * (float)W, fy = v * (float)H;
  unsigned x = (unsigned)fx, y = (unsigned)fy;
  float x1 = fx - (float)x, y1 = fy - (float)y;
  if (x > (unsigned)(W - 2)) x = (unsigned)(W - 2);
  if (y > (unsigned)(H - 2)) y = (unsigned)(H - 2);
  float x2 = 1.0f - x1, y2 = 1.0f - y1;
#define TX(xx, yy) V3(s->env[3 * ((size_t)(yy) * W + (xx))], s->env[3 * ((size_t)(yy) * W + (xx)) + 1], s->env[3 * ((size_t)(yy) * W + (xx)) + 2])
  v3 c1 = TX(x, y), c2 = TX(x + 1, y), c3 = TX(x, y + 1), c4 = TX(x + 1, y + 1);
#undef TX
  v3 c12 = vadd(vscale(c1, x2), vscale(c2, x1));
  v3 c34 = vadd(vscale(c3, x2), vscale(c4, x1));
  v3 c = vadd(vscale(c12, y2), vscale(c34, y1));
  return vscale(c, s->env_scale);
}

/* uniformSampleSphere (utils/random.hpp:7-15) */
v3 or_uniform_sphere(float u1, float u2) {
  float z = 1.0f - 2.0f * u1;
  float r = sqrtf(fmaxf(0.f, 1.f - z * z));
  float phi = (float)(2.0f * M_PI * u2);
  return V3(r * cosf(phi), r * sinf(phi), z);
}

static int find_interval(const float *cdf, int n, float u) {
  int lo = 0, hi = n; /* largest i in [0, n-1] with cdf[i] <= u */
  while (lo + 1 < hi) {
    int mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid; else hi = mid;
  }
  return lo;
}

/* 2-D piecewise-constant sampling of the environment (new). */
v3 or_env_sample_dir(const or_scene *s, float u0, float u1, float *pdf, float *uv) {
  int W = s->env_w, H = s->env_h;
  if (!(s->env_integral > 0.f)) {
    *pdf = (float)(1.0 / (4.0 * M_PI));
    if (uv) uv[0] = -1.f, uv[1] = 0.f;
    return or_uniform_sphere(u0, u1);
  }
  int y = find_interval(s->env_mcdf, H, u1);
  float dy = (u1 - s->env_mcdf[y]);
  float wy = s->env_mcdf[y + 1] - s->env_mcdf[y];
  if (wy > 0.f) dy /= wy;
  const float *cdf = &s->env_cdf[(size_t)y * (W + 1)];
  int x = find_interval(cdf, W, u0);
  float dx = (u0 - cdf[x]);
  float wx = cdf[x + 1] - cdf[x];
  if (wx > 0.f) dx /= wx;
  float u = ((float)x + dx) / (float)W, v = ((float)y + dy) / (float)H;
  if (uv) uv[0] = u, uv[1] = v;
  float theta = v * (float)M_PI, phi = u * 2.f * (float)M_PI;
  float st = sinf(theta);
  v3 dir = V3(-st * cosf(phi), -cosf(theta), st * sinf(phi));
  float f = s->env_func[(size_t)y * W + x];
  *pdf = (st > 0.f) ? (float)(f / s->env_integral / (2.0 * M_PI * M_PI * st)) : 0.f;
  return dir;
}

float or_env_pdf_dir(const or_scene *s, v3 dir) {
  int W = s->env_w, H = s->env_h;
  if (!(s->env_integral > 0.f)) return (float)(1.0 / (4.0 * M_PI));
  v3 dn = vnorm(dir);
  float theta = acosf(clampf(-dn.y, -1.f, 1.f));
  float phi = (float)(atan2f(-dn.z, dn.x) + M_PI);
  float u = (float)(phi / (2.0 * M_PI)), v = (float)(theta / M_PI);
  int x = (int)(u * W), y = (int)(v * H);
  if (x > W - 1) x = W - 1;
  if (y > H - 1) y = H - 1;
  if (x < 0) x = 0;
  if (y < 0) y = 0;
  float st = sinf(theta);
  if (!(st > 0.f)) return 0.f;
  return (float)(s->env_func[(size_t)y * W + x] / s->env_integral / (2.0 * M_PI * M_PI * st));
}

float or_env_pdf(const or_scene *s, const float *dir) { return or_env_pdf_dir(s, V3(dir[0], dir[1], dir[2])); }

int or_env_sample_batch(const or_scene *s, const float *u2, size_t n, float *dir, float *pdf) {
  for (size_t i = 0; i < n; ++i) {
    v3 d = or_env_sample_dir(s, u2[2 * i], u2[2 * i + 1], &pdf[i], NULL);
    dir[3 * i] = d.x; dir[3 * i + 1] = d.y; dir[3 * i + 2] = d.z;
  }
  return 0;
}

/* Hittable::pdf_value of the shape behind an area light, from o towards the
 * unit direction wi: AARectangle (primitive.hpp:244-256) d^2/(|cos| A) with
 * the hit's t, Sphere (primitive.hpp:172-184) 1/(2 pi (1 - cos theta_max));
 * 0 when the ray misses the shape (default tmin 1e-4, tmax 1e4). */
float or_shape_pdf(const or_scene *s, const or_light *L, v3 o, v3 wi) {
  float t;
  if (L->type == OR_LIGHT_QUAD) {
    const or_rect *r = &s->rects[L->src];
    if (!rect_hit(r, o, wi, 1e-4f, 1e4f, &t)) return 0.f;
    float area = (r->d0max - r->d0min) * (r->d1max - r->d1min);
    float d2 = t * t * vdot(wi, wi);
    v3 n = r->axis == 0 ? V3(1, 0, 0) : (r->axis == 1 ? V3(0, 1, 0) : V3(0, 0, 1));
    float cosine = fabsf(vdot(wi, n) / vlen(wi));
    return d2 / (cosine * area);
  }
  if (L->type == OR_LIGHT_SPHERE) {
    const or_sphere *sp = &s->spheres[L->src];
    if (!sphere_hit(sp, o, wi, 1e-4f, 1e4f, &t)) return 0.f;
    v3 oc = vsub(sp->c, o);
    float ctm = sqrtf(1.f - sp->r * sp->r / vdot(oc, oc));
    return (float)(1.0 / (2.0 * M_PI * (double)(1.f - ctm)));
  }
  return 0.f;
}

/* Hittable::get_sample of that shape from o (not normalised): a uniform
 * point of the rect minus o (primitive.cpp:125-146, C_rand(a, b) = a + (b-a) u),
 * or a direction in the cone the sphere subtends (primitive.hpp:186-192,
 * randomToSphere random.hpp:36-44, OrthoBasis around c - o). */
v3 or_shape_sample(const or_scene *s, const or_light *L, v3 o, float u, float v) {
  if (L->type == OR_LIGHT_QUAD) {
    const or_rect *r = &s->rects[L->src];
    float a = r->d0min + (r->d0max - r->d0min) * u;
    float b = r->d1min + (r->d1max - r->d1min) * v;
    v3 pt = r->axis == 0 ? V3(r->k, a, b) : (r->axis == 1 ? V3(a, r->k, b) : V3(a, b, r->k));
    return vsub(pt, o);
  }
  const or_sphere *sp = &s->spheres[L->src];
  v3 dir = vsub(sp->c, o);
  float d2 = vdot(dir, dir);
  float z = 1.f + v * (sqrtf(1.f - sp->r * sp->r / d2) - 1.f);
  float phi = 2.f * (float)M_PI * u;
  float x = cosf(phi) * sqrtf(1.f - z * z);
  float y = sinf(phi) * sqrtf(1.f - z * z);
  onb_t b = onb_make(vnorm(dir));
  return onb_to_world(&b, V3(x, y, z));
}

/* Light::get_sample / get_normal_at / L_emit (light/light.hpp:131-165,
 * 254-303; envLight.hpp:20-37,51-54,79-91).  Returns the sampled point. */
v3 or_light_point(const or_scene *s, const or_light *L, float u, float v, v3 *nrm) {
  if (L->type == OR_LIGHT_QUAD) {
    v3 e1 = vsub(L->v[1], L->v[0]), e2 = vsub(L->v[3], L->v[0]);
    v3 p = vadd(vadd(L->v[0], vscale(e1, u)), vscale(e2, v));
    *nrm = vnorm(vcross(e1, e2));
    return p;
  }
  if (L->type == OR_LIGHT_SPHERE) {
    float theta = (float)(2 * M_PI * u);
    float phi = acosf(2 * v - 1);
    v3 dir = V3(sinf(phi) * cosf(theta), sinf(phi) * sinf(theta), cosf(phi));
    v3 p = vadd(L->c, vscale(dir, L->r));
    *nrm = vnorm(vsub(p, L->c));
    return p;
  }
  (void)s;
  v3 d = or_uniform_sphere(u, v);
  v3 p = vscale(d, 1e5f);
  *nrm = vneg(vnorm(p));
  return p;
}

int or_light_sample_batch(const or_scene *s, uint32_t light, const float *u2, const float *ref,
                          size_t n, float *p, float *nrm, float *pdf, float *le) {
  if ((int)light >= s->n_lights) return -1;
  const or_light *L = &s->lights[light];
  for (size_t i = 0; i < n; ++i) {
    v3 nl;
    v3 pt = or_light_point(s, L, u2[2 * i], u2[2 * i + 1], &nl);
    v3 o = V3(ref[3 * i], ref[3 * i + 1], ref[3 * i + 2]);
    v3 w = vnorm(vsub(pt, o));
    float pd;
    v3 e;
    if (L->type == OR_LIGHT_ENV) {
      pd = (float)(1.0 / (4.0 * M_PI));
      e = or_env_lookup(s, w);
    } else if (L->type == OR_LIGHT_QUAD) {
      float d2 = vdot(vsub(pt, o), vsub(pt, o));
      float c = fabsf(vdot(L->n, vneg(w)));
      pd = d2 / (L->area * c);
      e = (vdot(nl, vneg(w)) > 0.f) ? L->le : vs(0.f);
    } else {
      v3 wc = vsub(L->c, o);
      float dc2 = vdot(wc, wc), r2 = L->r * L->r;
      if (dc2 <= r2) pd = (float)(1.f / (4.f * M_PI));
      else pd = (float)(1.f / (2.f * M_PI * (1.f - sqrtf(1.f - r2 / dc2))));
      e = (vdot(nl, vneg(w)) > 0.f) ? L->le : vs(0.f);
    }
    p[3 * i] = pt.x; p[3 * i + 1] = pt.y; p[3 * i + 2] = pt.z;
    nrm[3 * i] = nl.x; nrm[3 * i + 1] = nl.y; nrm[3 * i + 2] = nl.z;
    pdf[i] = pd;
    le[3 * i] = e.x; le[3 * i + 1] = e.y; le[3 * i + 2] = e.z;
  }
  return 0;
}
