/*
 * oracle.h -- CPU restatement of the reference renderer's per-sample hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this library is the checker for the HIP path
 * tracer in pathtracer_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never calls it.
 *
 * Every function cites the reference file:line it restates.  Where a
 * function is pinned only partly (no reference run is possible here, see
 * DESIGN.md), its comment says "parity unpinned" for the part left open.
 *
 * Data formats (materials, camera, render configuration) are the public
 * structs of include/pathtracer_amd.h, so a scene is described once and fed
 * to both the oracle and the product.
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#include "../include/pathtracer_amd.h"
#include "or_math.h"

#ifdef __cplusplus
extern "C" {
#endif

/* pt_render_cfg.flags bit read by the oracle only: the reference estimator
 * traverses each continuation ray twice, once for its emitter check and once
 * more in the recursion (maxdepth_nee.hpp:238-255).  The result is the same;
 * with this bit the oracle repeats the second traversal too, so that its
 * timing counts the reference's traversals (the CPU baseline).            */
#define OR_RENDER_REFERENCE_TRAVERSALS (1u << 24)

/* ---- samplers (or_sampler.c) ---- */
uint64_t or_mix_bits(uint64_t v);
uint32_t or_permutation_element(uint32_t i, uint32_t l, uint32_t p);
uint64_t or_murmur64a(const unsigned char *key, size_t len, uint64_t seed);
uint64_t or_hash_base_digit_seed(int32_t base, int32_t digit_index, uint32_t seed);
float or_radical_inverse(int base_index, uint64_t a);
double or_owen_radical_inverse(int base_index, uint64_t a, uint32_t hash);
double or_scrambled_radical_inverse(int base_index, uint64_t a, uint32_t seed);
float or_halton_dim(uint32_t strategy, uint32_t seed, uint32_t dim, uint64_t index);
uint32_t or_sobol_matrix_value(uint32_t dim, uint32_t index);
float or_sobol_owen(uint32_t index, uint32_t dim, uint32_t seed);
float or_stratified(uint32_t pixel_seed, uint32_t k, uint32_t spp_x, uint32_t spp_y,
                    uint32_t dim);
uint32_t or_hash32(uint32_t x);
uint32_t or_rng_u32(uint32_t a, uint32_t b, uint32_t c);
float or_rng(uint32_t a, uint32_t b, uint32_t c);
uint32_t or_pixel_seed(uint32_t pixel, uint32_t seed);
float or_sample_dim(uint32_t sampler, uint32_t seed, uint32_t spp_x, uint32_t spp_y,
                    uint32_t pixel, uint32_t k, uint64_t global_index, uint32_t dim);
int or_sample_batch(uint32_t sampler, uint32_t seed, uint32_t spp, const uint32_t *dim,
                    const uint64_t *index, const uint32_t *pixel, size_t n, float *out);

/* ---- materials (or_bsdf.c) ---- */
enum { OR_REFL = 1, OR_TRANS = 2, OR_DIFFUSE = 4, OR_GLOSSY = 8, OR_SPECULAR = 16 };
typedef struct {
  v3 wi, att;
  float pdf;
  uint32_t flags;
} or_srec;
typedef struct {
  float t;
  v3 p, n;      /* n opposes the incoming ray (HitRecord::set_normal) */
  v3 ng_light;  /* un-flipped geometric normal (for one-sided emitters) */
  v3 tangent;   /* HitRecord::hair_tangent (curves), else 0             */
  float u, v;   /* HitRecord::texCoords                                 */
  v3 tex;       /* a textured material's colour at (u, v)               */
  int has_tex;  /* the material is textured: tex replaces its colour    */
  int inside;   /* HitRecord::is_inside = front face                   */
  int prim;
  uint32_t mat;
} or_hit;
int or_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float uc, float u0,
               float u1, float aux, or_srec *s);
v3 or_bsdf_f(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h);
v3 or_cosine_dir(v3 n, float u0, float u1);
float or_cosine_pdf(v3 n, v3 wi);
float or_bsdf_pdf(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h);
v3 or_emit(const pt_material_desc *m, v3 ray_d, const or_hit *h);
int or_is_emitter(const pt_material_desc *m);
int or_bsdf_eval_batch(const pt_material_desc *m, const float *wo, const float *wi,
                       const float *nrm, const uint8_t *inside, size_t n, float *f,
                       float *pdf);
int or_bsdf_sample_batch(const pt_material_desc *m, const float *wo, const float *nrm,
                         const uint8_t *inside, const float *uc, const float *u2,
                         const float *aux, size_t n, float *wi, float *f, float *pdf,
                         uint32_t *flags, uint8_t *ok);

/* ---- scene, geometry, lights, integrator (or_scene.c, or_render.c) ---- */
typedef struct or_scene or_scene;
or_scene *or_scene_create(void);
void or_scene_destroy(or_scene *s);
int or_scene_add_material(or_scene *s, const pt_material_desc *d);
int or_scene_add_mesh(or_scene *s, const float *pos, uint32_t nv, const int32_t *idx,
                      uint32_t nt, uint32_t mat, uint32_t flags);
int or_scene_add_mesh_normals(or_scene *s, const float *pos, uint32_t nv, const int32_t *idx,
                              uint32_t nt, const float *nrm, uint32_t nn, const int32_t *nidx,
                              uint32_t mat, uint32_t flags);
int or_scene_add_sphere(or_scene *s, const float *c, float r, uint32_t mat, uint32_t flags);
int or_scene_add_rect(or_scene *s, int axis, float k, float d0min, float d0max, float d1min,
                      float d1max, uint32_t mat, uint32_t flags);
int or_scene_set_env(or_scene *s, const float *rgb, uint32_t w, uint32_t h, float scale);
int or_scene_add_curve(or_scene *s, const float *cps, float w0, float w1, uint32_t mat, uint32_t flags);
int or_scene_set_camera(or_scene *s, const pt_camera_desc *c);
/* HomogeneousMedium(sigma_a, sigma_s, sigma_scale, Le, Le_scale, g); sigma_a NULL removes it. */
int or_scene_set_medium(or_scene *s, const float *sigma_a, const float *sigma_s, float sigma_scale, float g);
int or_scene_add_texture_image(or_scene *s, const float *rgb, uint32_t w, uint32_t h);
int or_scene_add_texture_checker(or_scene *s, const float *c1, const float *c2, float scale);
int or_scene_set_mesh_uv(or_scene *s, uint32_t mesh, const float *uv, uint32_t nuv, const int32_t *uvidx,
                         uint32_t nt);
v3 or_texture(const or_scene *s, uint32_t id, float u, float v);
int or_texture_batch(const or_scene *s, uint32_t id, const float *uv, size_t n, float *out);
int or_hit_uv_batch(const or_scene *s, const float *rays, size_t n, float *uv);
int or_scene_commit(or_scene *s);
int or_scene_num_lights(const or_scene *s);
int or_surface_batch(const or_scene *s, const float *rays, size_t n, float *nrm, int32_t *inside);
int or_intersect_batch(const or_scene *s, const float *rays, size_t n, int any_hit,
                       int32_t *prim, float *t);
int or_brute_intersect_batch(const or_scene *s, const float *rays, size_t n,
                             int32_t *prim, float *t);
int or_light_sample_batch(const or_scene *s, uint32_t light, const float *u2,
                          const float *ref, size_t n, float *p, float *nrm, float *pdf,
                          float *le);
int or_env_sample_batch(const or_scene *s, const float *u2, size_t n, float *dir,
                        float *pdf);
float or_env_pdf(const or_scene *s, const float *dir);
int or_camera_rays(const or_scene *s, const pt_render_cfg *cfg, const uint32_t *pixel,
                   const uint32_t *k, size_t n, float *rays);
/* Renders cfg into out_rgb (W*H*3 per-pixel means).  Optional: out_var
 * (W*H*3 per-pixel sample variance), out_samples (W*H*spp*3 per-sample
 * radiance), counters (8 x uint64: paths, segments, shadow rays, nan
 * samples, ...).  threads <= 0: all OpenMP threads.                       */
int or_render(const or_scene *s, const pt_render_cfg *cfg, float *out_rgb, float *out_var,
              float *out_samples, uint64_t *counters, int threads);
/* ---- the guided estimator (or_render.c) ----
 * The device's learned-guiding estimator (no counterpart in the reference)
 * restated with a fixed vMF mixture in place of the network: every point of a
 * guided material kind at vertex d < verts samples the mixture with
 * probability c = mix[40], else the BSDF, and weights the direction by the
 * one-sample MIS pdf c p_guide + (1 - c) p_bsdf, which NEE's power heuristic
 * and the next vertex's emitter MIS also use.  mix is the processed mixture
 * (PT_GUIDE_MIX floats: per lobe (kappa mu, log(pi kappa / (2 pi (1 -
 * e^{-2 kappa}))) - kappa), the 8 weights pi, c).  cfg->estimator
 * OR_EST_GUIDED selects it in or_render.                                  */
#define OR_EST_GUIDED 6
int or_scene_set_guide(or_scene *s, const float *mix, uint32_t kinds, uint32_t verts);
float or_guide_pdf(const float *mix, v3 w);
v3 or_guide_sample(const float *mix, float u0, float u1, float u2);
int or_guide_eval_batch(const float *mix, const float *dirs, const float *u, size_t n, float *pdf,
                        float *sample_dir);
int or_guidable(const pt_material_desc *m, uint32_t kinds);
/* Renders only the path samples in [first, first+count) of the pixel-major
 * sample order and returns their summed luminance (for timing).          */
double or_render_range(const or_scene *s, const pt_render_cfg *cfg, uint64_t first,
                       uint64_t count, int threads, uint64_t *counters);

#ifdef __cplusplus
}
#endif
#endif
