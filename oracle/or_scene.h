/*
 * or_scene.h -- internal scene layout of the CPU oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#ifndef OR_SCENE_H
#define OR_SCENE_H
#include "oracle.h"

typedef struct { v3 p[3]; int mesh; int smooth; v3 n[3]; } or_tri; /* n: corner normals when smooth */
typedef struct {
  v3 bmin, bmax;
  int left, right;
  int leaf_first, leaf_count;
} or_mnode;
typedef struct {
  int first_tri, n_tris;
  uint32_t mat, flags, seq;
  or_mnode *nodes;
  int n_nodes, cap_nodes;
  int *leaf_ids;
  int n_leaf, cap_leaf;
  v3 box_lo, box_hi;  /* TriangleMesh::meshAABB (triangleMesh.cpp:10-18)   */
  float *uv;          /* MeshData::uvs (2 floats each), NULL: planar UVs   */
  int32_t *uvidx;     /* MeshData::uvIndices, 3 per triangle               */
} or_mesh;
/* ImageTexture texels (texel (x, y) at rgb[3 (y w + x)], y = 0 the bottom
 * row) or a CheckerTexture (texture/texture.hpp:49-94). */
typedef struct {
  int kind; /* 0 image, 1 checker */
  int w, h;
  float *rgb;
  v3 c1, c2;
  float scale;
} or_tex;
typedef struct { v3 c; float r; uint32_t mat, flags, seq; int light; } or_sphere;
typedef struct {
  int axis;
  float k, d0min, d0max, d1min, d1max;
  uint32_t mat, flags, seq;
  int light;
} or_rect;
/* Curve (primitives/curve.hpp:13-140) with its 9 Bezier points and
 * derivatives precomputed as Curve::intersect computes them. */
typedef struct {
  v3 cps[4];
  float w0, w1;
  v3 P[9], dP[9];
  uint32_t mat, flags;
} or_curve;
enum { OR_LIGHT_QUAD = 1, OR_LIGHT_SPHERE = 2, OR_LIGHT_ENV = 3 };
typedef struct {
  int type;
  v3 v[4], n; /* quad */
  v3 c;       /* sphere */
  float r, area;
  v3 le;      /* texture colour * intensity */
  int src;    /* the rect (quad) or sphere it was discovered from */
} or_light;

struct or_scene {
  pt_material_desc *mats;
  int n_mats, cap_mats;
  or_tri *tris;
  int n_tris, cap_tris;
  or_mesh *meshes;
  int n_meshes, cap_meshes;
  or_sphere *spheres;
  int n_spheres, cap_spheres;
  or_rect *rects;
  int n_rects, cap_rects;
  or_curve *curves;
  int n_curves, cap_curves;
  or_tex *tex;
  int n_tex, cap_tex;
  or_light *lights;
  int n_lights;
  uint32_t seq;
  float *env;
  int env_w, env_h;
  float env_scale;
  float *env_func, *env_cdf, *env_marg, *env_mcdf;
  float env_integral;
  pt_camera_desc cam;
  int has_cam;
  /* SceneInfo::global_medium: HomogeneousMedium (homogeneous.hpp:31-68) */
  v3 med_sa, med_ss;
  float med_g;
  int has_medium;
  v3 cam_h, cam_v, cam_horiz, cam_vert, cam_bl_rel;
  int committed;
  /* the guided estimator's mixture (one for every guided point), the
   * material kinds it guides and how many vertices (or_scene_set_guide) */
  float gmix[48];
  uint32_t gkinds, gverts;
};

int or_scene_intersect(const or_scene *s, v3 o, v3 d, float tmin, float tmax, int any,
                       or_hit *h, uint64_t *cnt);
int or_prim_light(const or_scene *s, int prim);
v3 or_env_lookup(const or_scene *s, v3 dir);
v3 or_env_lookup_uv(const or_scene *s, float u, float v);
v3 or_uniform_sphere(float u1, float u2);
/* uv (may be NULL): the sampled texture coordinates, uv[0] < 0 for uniform sampling */
v3 or_env_sample_dir(const or_scene *s, float u0, float u1, float *pdf, float *uv);
float or_env_pdf_dir(const or_scene *s, v3 dir);
v3 or_light_point(const or_scene *s, const or_light *L, float u, float v, v3 *nrm);
float or_shape_pdf(const or_scene *s, const or_light *L, v3 o, v3 wi);
v3 or_shape_sample(const or_scene *s, const or_light *L, v3 o, float u, float v);
#endif
