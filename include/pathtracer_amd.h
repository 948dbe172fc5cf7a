/*
 * pathtracer_amd.h -- C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * The reference renderer (747745124/PathTracer) has no FFI layer: its
 * boundary is a set of C++ abstract classes bound at compile time.  Each
 * entry point below replaces one of those interfaces; the comment on each
 * names the reference file:line it stands in for.  A source-compatible C++
 * scene API that lowers onto these calls lives in
 * include/pathtracer_amd/scene.hpp.
 *
 * Conventions
 *   - Every call returns a pt_status (0 = PT_OK).  Nothing throws across the
 *     ABI; pt_last_error() returns a thread-local message for the last
 *     failing call on this thread.
 *   - Host buffers are caller-owned.  Device buffers are owned by pt_scene.
 *   - All arrays are plain C arrays; vectors are packed float triples.
 *   - A pt_scene is used by one host thread at a time (thread-compatible,
 *     not thread-safe), and owns one HIP stream.
 */
#ifndef PATHTRACER_AMD_H
#define PATHTRACER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pt_status;
enum {
  PT_OK = 0,
  PT_ERR_INVALID = 1,     /* bad argument / invalid scene state            */
  PT_ERR_DEVICE = 2,      /* HIP runtime error                             */
  PT_ERR_NOMEM = 3,       /* host or device allocation failed              */
  PT_ERR_UNSUPPORTED = 4, /* feature not available                         */
  PT_ERR_NO_DEVICE = 5    /* no gfx950 device visible                      */
};

/* ------------------------------------------------------------------ */
/* Materials  (reference: class Material, material/material.hpp:15-77) */
/* ------------------------------------------------------------------ */
enum pt_mat_kind {
  PT_MAT_LAMBERTIAN = 1,     /* material.hpp:79-137        p: albedo[3]          */
  PT_MAT_EMITTER = 2,        /* material.hpp:467-506       p: color[3], intensity */
  PT_MAT_CONDUCTOR = 3,      /* conductor.hpp:4-145        p: eta[3], k[3], ax, ay */
  PT_MAT_MFDIELECTRIC = 4,   /* dielectric.hpp:5-167       p: eta, ax, ay        */
  PT_MAT_THINDIELECTRIC = 5, /* thinDielectric.hpp:3-70    p: eta, split(0/1)    */
  PT_MAT_PHONG = 6,          /* material.hpp:303-465       p: diffuse[3], specular[3], ambient[3], shininess */
  PT_MAT_DISNEY = 7,         /* disneyPrincipledBSDF.hpp:8-99 p: baseColor[3], specTrans, metallic,
                                subsurface, specular, roughness, specularTint, anisotropic, sheen,
                                sheenTint, clearcoat, clearcoatGloss, eta   (constructor order) */
  PT_MAT_HAIR = 8,           /* hairMarsch.hpp:3-158       p: sigma_a[3], h, eta, beta_m, beta_n, alpha */
  PT_MAT_KAJIYA_KAY = 9      /* no reference code (README.md:74 lists it)  p: diffuse[3], specular[3], exponent */
};

#define PT_MAT_NPARAMS 30
typedef struct pt_material_desc {
  uint32_t kind;              /* pt_mat_kind                                */
  uint32_t flags;             /* 0, or PT_MAT_TEXTURE(id)                    */
  float p[PT_MAT_NPARAMS];    /* kind-specific parameters, see pt_mat_kind  */
} pt_material_desc;           /* 128 bytes                                  */
/* pt_material_desc.flags: the material's colour -- Lambertian albedo
 * (material.hpp:83-100) or Disney baseColor (disneyPrincipledBSDF.hpp) --
 * is texture `id` (pt_scene_add_texture_*) at the hit's texture coordinates
 * (HitRecord::texCoords) instead of p[0..2].  Other kinds refuse it.        */
#define PT_MAT_TEXTURE(id) (0x1u | ((uint32_t)(id) << 8))

/* Primitive flags.  The reference's wrappers change HitRecord::is_inside:
 * FlipFace toggles it (objectTransform.hpp:21); Translate/Rotate/Scale re-run
 * set_normal on an already ray-opposed normal, which forces it true
 * (objectTransform.hpp:60,156; objectTransform.cpp:132,172,211).  Applied in
 * this order: FORCE_INSIDE sets it true, then FLIP_FACE toggles it, so a
 * FlipFace around a transformed object (both flags) yields false.
 * NOT_LIGHT keeps an emissive rect/sphere out of light discovery, which in
 * the reference only looks at top-level objects (lightDiscovery.hpp:10-60):
 * emitters inside a Box or a nested list are hit but never sampled. */
enum {
  PT_PRIM_FLIP_FACE = 1u << 0,
  PT_PRIM_FORCE_INSIDE = 1u << 1,
  PT_PRIM_NOT_LIGHT = 1u << 2
};

/* ------------------------------------------------------------------ */
/* Camera  (reference: PerspectiveCamera, base/camera.cpp:7-56)        */
/* ------------------------------------------------------------------ */
typedef struct pt_camera_desc {
  float fov;          /* vertical field of view, radians (camera.cpp:12)        */
  float aspect;       /* width / height                                         */
  float focal_dist;   /* image-plane distance                                   */
  float aperture;     /* reference semantics: lens half-width = focal_dist/aperture/2,
                         lens point uniform on a square (camera.cpp:44-50);
                         0 or +inf = pinhole                                     */
  float up[3];
  float front[3];
  float pos[3];
  float pos1[3];      /* camera position at shutter close (motion blur)         */
  float shutter;      /* 0 = static camera; >0: position = lerp(pos,pos1,time)   */
  uint32_t normalize_dir; /* 1: camera rays carry unit directions               */
} pt_camera_desc;

/* ------------------------------------------------------------------ */
/* Samplers  (reference: HaltonSampler, sampler/sampler.hpp:16-83)     */
/* ------------------------------------------------------------------ */
enum pt_sampler_kind {
  PT_SAMPLER_HALTON_OWEN = 0,  /* RandomStrategy::Owen  (default, sampler.hpp:67-71)   */
  PT_SAMPLER_HALTON_PERM = 1,  /* RandomStrategy::PermuteDigits (sampler.hpp:60-65)     */
  PT_SAMPLER_HALTON_NONE = 2,  /* RandomStrategy::None  (sampler.hpp:57-58)            */
  PT_SAMPLER_SOBOL = 3,        /* Owen-scrambled Sobol; no reference code (README.md:98) */
  PT_SAMPLER_STRATIFIED = 4    /* per-pixel jittered strata with permuted strata order  */
};

/* Integrator estimators (reference: getRayColor, integrator/maxdepth_nee.hpp:9-256) */
enum pt_estimator {
  PT_EST_REFERENCE_NEE = 0, /* the reference's estimator: N light samples per
                               non-specular vertex, emitter hits excluded after
                               non-specular vertices, env radiance also taken on
                               BSDF misses (maxdepth_nee.hpp:26-37,91-255)       */
  PT_EST_NEE_MIS = 1,       /* one light sample + BSDF sample under the power
                               heuristic; environment sampled by its 2-D CDF     */
  PT_EST_NAIVE = 2,         /* maxdepth_naive.hpp:9-75: no NEE; each bounce samples
                               a uniform mixture of light_samples shape pdfs and
                               the BSDF; misses and the depth limit return bg     */
  PT_EST_ROULETTE = 3,      /* roulette_naive.hpp:9-78: the same mixture with
                               Russian roulette q = clamp(max(f), 0.5, 1) at every
                               vertex; the depth limit returns 0                  */
  PT_EST_RESERVOIR = 4,     /* maxdepth_reservoir_di.hpp:9-131: direct light
                               resampled from light_samples shadowed candidates;
                               a vertex whose continuation hits an emitter adds 0 */
  PT_EST_VOLUME = 5         /* maxdepth_volume.hpp:72-118: single scattering in
                               the scene's homogeneous medium (pt_scene_set_medium):
                               a free-flight distance, light_samples shadowed light
                               samples with HG phase and transmittance; a surface
                               reached first adds only its emission               */
};

typedef struct pt_render_cfg {
  uint32_t width, height;
  uint32_t spp_x, spp_y;   /* samples per pixel = spp_x*spp_y (framebuffer.cpp:131-145) */
  uint32_t max_depth;      /* getRayColor(max_depth): vertices 0..max_depth-1 shade  */
  uint32_t sampler;        /* pt_sampler_kind                                        */
  uint32_t estimator;      /* pt_estimator                                           */
  uint32_t rr_depth;       /* first vertex index with Russian roulette; 0 = off      */
  uint32_t light_samples;  /* NEE samples per vertex for PT_EST_REFERENCE_NEE (4 = 2x2) */
  uint32_t seed;           /* seeds the counter-based random stream                   */
  float bg[3];             /* background radiance when there is no environment        */
  uint32_t spp_begin;      /* render samples [spp_begin, spp_begin+spp_count)         */
  uint32_t spp_count;      /* 0 = all samples                                         */
  uint32_t wavefront;      /* path slots in flight; 0 = automatic                     */
  uint32_t flags;          /* PT_RENDER_* bits                                        */
} pt_render_cfg;

enum {
  PT_RENDER_COUNTERS = 1u << 0,   /* collect traversal counters into pt_stats       */
  PT_RENDER_SORT_HITS = 1u << 1,  /* order each 256-entry tile of hits by material
                                     kind (misses first) before shading, so a wave
                                     shades one kind where the tile allows.  Off by
                                     default: it did not pay on the measured scenes */
  PT_RENDER_ACCUMULATE = 1u << 2, /* add into the film instead of clearing it        */
  PT_RENDER_KEEP_SAMPLES = 1u << 3, /* keep per-sample radiance (pt_film_read_samples) */
  PT_RENDER_TIMING = 1u << 4      /* HIP-event timing of every kernel launch        */
};

typedef struct pt_stats {
  uint64_t paths;          /* path samples traced                                */
  uint64_t segments;       /* extension (closest-hit) rays                        */
  uint64_t shadow_rays;    /* shadow (any-hit) rays                               */
  uint64_t node_visits;    /* BVH node-pair fetches, extension rays               */
  uint64_t prim_tests;     /* primitive tests, extension rays                     */
  uint64_t shadow_node_visits;
  uint64_t shadow_prim_tests;
  uint64_t nan_samples;    /* path samples discarded for NaN/Inf radiance        */
  uint64_t extend_launches;
  double ms_total;         /* wall time of the render call (device work)         */
  double ms_generate, ms_extend, ms_shade, ms_shadow, ms_film; /* HIP-event sums   */
} pt_stats;

/* ------------------------------------------------------------------ */
/* Scene  (reference: SceneInfo, render/renderManager.hpp:35-158;       */
/*         ObjectList::addObject objectList.hpp:73)                     */
/* ------------------------------------------------------------------ */
typedef struct pt_scene pt_scene;

pt_status pt_scene_create(pt_scene **out);
void pt_scene_destroy(pt_scene *scene);

/* Material constructors (material/ headers).  Returns the material id.        */
pt_status pt_scene_add_material(pt_scene *scene, const pt_material_desc *desc,
                                uint32_t *out_id);
/* TriangleMesh(MeshData, material)  (primitives/triangleMesh.cpp:5-26).
 * pos: nverts*3 floats; idx: ntris*3 int32 (MeshData::indices).  Without
 * per-vertex normals shading uses the geometric normal, meshBVH.cpp:121.    */
pt_status pt_scene_add_mesh(pt_scene *scene, const float *pos, uint32_t nverts,
                            const int32_t *idx, uint32_t ntris,
                            uint32_t material, uint32_t flags);
/* TriangleMesh with per-corner shading normals (MeshData::normals and
 * normalIndices, mesh_io/meshData.hpp:6-13): the hit interpolates
 * (1-u-v) n0 + u n1 + v n2 with the hit's barycentrics, normalises it and
 * turns it to the geometric normal's side (meshBVH.cpp:116-130).
 * nrm: nnormals*3 floats; nidx: ntris*3 int32.                             */
pt_status pt_scene_add_mesh_normals(pt_scene *scene, const float *pos, uint32_t nverts,
                                    const int32_t *idx, uint32_t ntris,
                                    const float *nrm, uint32_t nnormals,
                                    const int32_t *nidx, uint32_t material,
                                    uint32_t flags);
/* Replaces the vertex positions of mesh `mesh` (meshes are numbered in the
 * order they were added, from 0) with pos (nverts*3 floats, the vertex count
 * it was added with; its index buffer and shading normals are kept).  Takes
 * effect at the next pt_scene_commit.  No reference counterpart: the
 * reference builds its scene once per render (render/renderManager.hpp:80-81);
 * this is for geometry that moves between frames, with PT_BUILD_DEVICE_LBVH. */
pt_status pt_scene_update_mesh(pt_scene *scene, uint32_t mesh, const float *pos, uint32_t nverts);
/* Sphere(center, radius, material)  (base/primitive.hpp:81-203)           */
pt_status pt_scene_add_sphere(pt_scene *scene, const float center[3],
                              float radius, uint32_t material, uint32_t flags);
/* AARectangle<Axis>(k, d0min, d0max, d1min, d1max, material)
 * (base/primitive.hpp:205-273).  axis: 0=X, 1=Y, 2=Z.                      */
pt_status pt_scene_add_rect(pt_scene *scene, int axis, float k, float d0min,
                            float d0max, float d1min, float d1max,
                            uint32_t material, uint32_t flags);
/* Curve(cps, width0, width1, type, material)  (primitives/curve.hpp:13-140):
 * a cubic Bezier (cps: 4 control points, 12 floats) hit as 8 cylinder
 * segments, radius the mean of the widths interpolated at each segment's
 * ends.  The hit sets the hair tangent (HitRecord::hair_tangent) that
 * HairMarschner::scatter frames on.  Every type intersects as a cylinder, as
 * in the reference.  Curves are never lights.  Their primitive ids follow the
 * rects', 8 per curve in insertion order.                                   */
enum { PT_CURVE_FLAT = 0, PT_CURVE_CYLINDER = 1, PT_CURVE_RIBBON = 2 };
pt_status pt_scene_add_curve(pt_scene *scene, const float cps[12], float width0,
                             float width1, uint32_t type, uint32_t material,
                             uint32_t flags);
/* Textures (texture/texture.hpp:20-94), referenced by materials through
 * PT_MAT_TEXTURE(id); ids count from 0 in the order of adding.
 * ImageTexture(file, flip_y, isHDR) after loading: w*h RGB texels, texel
 * (x, y) at rgb[(y*w + x)*3], y = 0 the image's bottom row (texture.cpp:
 * 21-45); looked up bilinearly with the reference's clamp of the lower
 * texel to w-2, h-2 (texture.cpp:111-142), so w, h >= 2.                   */
pt_status pt_scene_add_texture_image(pt_scene *scene, const float *rgb, uint32_t w, uint32_t h,
                                     uint32_t *out_id);
/* CheckerTexture(color1, color2, scale) (texture.cpp:151-183): color1 where
 * sin(u scale) sin(v scale) < 0, else color2.                              */
pt_status pt_scene_add_texture_checker(pt_scene *scene, const float color1[3], const float color2[3],
                                       float scale, uint32_t *out_id);
/* Texture coordinates of mesh `mesh` (meshes numbered in adding order from
 * 0): MeshData::uvs / uvIndices (mesh_io/meshData.hpp:6-13), uv: nuv*2
 * floats, uvidx: 3 int32 per triangle.  A hit interpolates (1-u-v) uv0 +
 * u uv1 + v uv2 with its barycentrics (meshBVH.cpp:133-138).  A mesh without
 * them takes TriangleMesh's planar fallback (triangleMesh.cpp:38-51).      */
pt_status pt_scene_set_mesh_uv(pt_scene *scene, uint32_t mesh, const float *uv, uint32_t nuv,
                               const int32_t *uvidx, uint32_t ntris);
/* EnvironmentLight(hdri, scale)  (light/envLight.hpp:4-92).  rgb holds
 * w*h texels, texel (x,y) at rgb[(y*w+x)*3], y=0 being the direction -Y
 * (the reference stores texels[x][y] after a vertical flip, texture.cpp:39-42). */
pt_status pt_scene_set_env(pt_scene *scene, const float *rgb, uint32_t w,
                           uint32_t h, float scale);
pt_status pt_scene_set_camera(pt_scene *scene, const pt_camera_desc *cam);
/* HomogeneousMedium(sigma_a, sigma_s, sigma_scale, Le, Le_scale, g) as
 * SceneInfo::global_medium (medium/homogeneous.hpp:31-68,
 * renderManager.hpp:48): the scene-filling medium PT_EST_VOLUME renders.
 * sigma_a == NULL removes it.  Le is kept but, as in the reference's
 * integrator, not rendered.                                                */
pt_status pt_scene_set_medium(pt_scene *scene, const float sigma_a[3], const float sigma_s[3],
                              float sigma_scale, const float le[3], float le_scale, float g);

typedef struct pt_build_cfg {
  uint32_t leaf_size;   /* max primitives per BVH leaf (0 = 4)                */
  uint32_t sah_bins;    /* SAH bins per axis (0 = 16)                         */
  uint32_t lds_levels;  /* top levels of 4-wide BVH nodes traversal keeps in
                           LDS, whole levels up to 21 nodes (0 = 3)          */
  uint32_t flags;       /* PT_BUILD_*                                         */
} pt_build_cfg;

/* Builds the BVH on the device instead of the host: Morton order of the
 * primitive centroids, the binary radix tree of Karras (HPG 2012), a refit,
 * and the same 4-wide node layout as the host SAH build (sah_bins unused).
 * Replaces the host builds BVHNode (primitives/bvh.hpp:46-102) and
 * MeshBVHNode (primitives/meshBVH.cpp:6-53) for scenes rebuilt every frame:
 * milliseconds to build, some slower to trace than the SAH tree.  Hits are
 * the same closest hits either way.                                        */
#define PT_BUILD_DEVICE_LBVH 1u
/* Traversal reads every BVH node from global memory (no LDS copy of the top
 * levels).                                                                  */
#define PT_BUILD_NO_LDS_NODES 2u
/* Node format of any-hit (shadow-ray) traversal.  Every commit builds the
 * 128-byte nodes (four full-precision child boxes) that closest-hit
 * traversal reads, and a 64-byte copy whose child boxes are 8-bit offsets
 * from a per-node origin with per-axis power-of-two scales, rounded outwards
 * (conservative: the same hits).  Any hit reads the 64-byte copy when the
 * 128-byte nodes exceed 4 MiB (about 32k nodes), where it is bound by the
 * node bytes it fetches; these flags force one format.  A batch's shadow rays
 * read the 64-byte nodes when any of its scenes selects them.               */
#define PT_BUILD_ANY_HIT_QNODES 4u
#define PT_BUILD_ANY_HIT_FULL_NODES 8u

/* Builds the BVH, discovers emissive rects/spheres as lights
 * (base/lightDiscovery.hpp:5-81) and uploads everything to the device.      */
pt_status pt_scene_commit(pt_scene *scene, const pt_build_cfg *cfg);

typedef struct pt_scene_info {
  uint32_t n_materials, n_triangles, n_spheres, n_rects, n_lights;
  uint32_t bvh_nodes, bvh_depth;
  uint64_t device_bytes;
  double build_ms;  /* the whole commit                                       */
  double bvh_ms;    /* the BVH build alone (with PT_BUILD_DEVICE_LBVH: the
                       records' upload and the device kernels, waited for)    */
  uint32_t flags;   /* PT_INFO_*                                              */
} pt_scene_info;
/* pt_scene_info.flags: any-hit traversal reads the 64-byte quantised nodes
 * (chosen for trees above 4 MiB, or forced by PT_BUILD_ANY_HIT_QNODES).     */
#define PT_INFO_ANY_HIT_QNODES 1u
pt_status pt_scene_get_info(const pt_scene *scene, pt_scene_info *info);

/* SceneInfo::render (renderManager.hpp:51-158): renders into the scene's
 * device film.  The film holds per-pixel means, row j=0 at the bottom
 * (FrameBuffer::setPixelColor(j,i), renderManager.hpp:149).                */
pt_status pt_render(pt_scene *scene, const pt_render_cfg *cfg, pt_stats *stats);
/* Copies the film (width*height*3 floats, row-major, j=0 first) to host.   */
pt_status pt_film_read(pt_scene *scene, float *out_rgb);
/* Per-sample radiance of the last render when PT_RENDER keeps it
 * (n = width*height*spp*3 floats), ordered pixel-major then sample.        */
pt_status pt_film_read_samples(pt_scene *scene, float *out_rgb, size_t n);

/* pt_render without any device allocation.  No counterpart in the
 * reference (whose renderer allocates per frame, renderManager.hpp:80-95):
 * it is for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * scene must be committed (the commit allocates its device data).  The film
 * is the caller's device buffer of width*height*3 floats; it receives the
 * per-pixel means, row-major with j = 0 first (what pt_film_read returns
 * after pt_render).  The wavefront, pixel offsets and counters live in one
 * static 64 MiB device array per GPU, held until the process ends: at most
 * 131,072 path slots in flight (cfg->wavefront 0 means that maximum; larger
 * values are refused) and 65,536 samples per pixel.  Not supported (refused
 * with PT_ERR_UNSUPPORTED): PT_RENDER_KEEP_SAMPLES, PT_RENDER_TIMING,
 * PT_RENDER_ACCUMULATE, and the permuted-digit Halton sampler for a seed
 * whose tables the scene has not built.  The scene's device must be
 * current.  Calls on one GPU share the static array, so calls that overlap
 * on one GPU (from several threads or streams) race.                       */
pt_status pt_render_noalloc(pt_scene *scene, const pt_render_cfg *cfg, float *film, pt_stats *stats);

/* ------------------------------------------------------------------ */
/* Launch tuning and the autotuner (no counterpart in the reference,    */
/* whose OpenMP loop has no launch parameters, renderManager.hpp:99).   */
/* ------------------------------------------------------------------ */
typedef struct pt_tuning {
  uint32_t wavefront;            /* path slots in flight (0 = 64M); a render's
                                    pt_render_cfg.wavefront overrides it          */
  uint32_t trav_blocks_per_cu;   /* traversal grid, blocks of 256 per CU (0 = 8) */
  uint32_t shade_blocks_per_cu;  /* shading grid cap, blocks per CU (0 = 32)     */
  uint32_t leaf_size;            /* BVH leaf size of the last commit (autotuner
                                    output; pt_scene_set_tuning ignores it)     */
  uint32_t sort_bits;            /* bits of the material key shading orders its
                                    tiles by: 0 = queue order, 4 = material kind
                                    (as PT_RENDER_SORT_HITS, which forces it)   */
} pt_tuning;
pt_status pt_scene_set_tuning(pt_scene *scene, const pt_tuning *tuning);

enum {
  PT_TUNE_WAVEFRONT = 1u << 0,   /* try wavefront sizes 4M..64M                  */
  PT_TUNE_GRIDS = 1u << 1,       /* try traversal / shading grid sizes           */
  PT_TUNE_LEAF = 1u << 2,        /* rebuild the BVH with leaf sizes 1, 2, 4, 8   */
  PT_TUNE_SORT = 1u << 3         /* try sort-key bits 0 and 4                     */
};
typedef struct pt_autotune_cfg {
  uint32_t flags;      /* PT_TUNE_* (0 = all)                                     */
  uint32_t spp_div;    /* time candidates on 1/spp_div of the samples (0 = 8)     */
  uint32_t repeats;    /* timed renders per candidate, the fastest kept (0 = 1)  */
} pt_autotune_cfg;
/* Times short renders of cfg's frame for each candidate and keeps the
 * fastest: the result is stored with pt_scene_set_tuning, and with
 * PT_TUNE_LEAF the scene stays committed with the best leaf size.          */
pt_status pt_autotune(pt_scene *scene, const pt_render_cfg *cfg,
                      const pt_autotune_cfg *tc, pt_tuning *out);

/* ------------------------------------------------------------------ */
/* Batched rendering (no counterpart in the reference: it renders one   */
/* SceneInfo per call, renderManager.hpp:51-158).  Many renders of      */
/* uneven resolution, spp and scene size share one wavefront: their     */
/* path samples are concatenated and cut into chunks, and one captured  */
/* HIP graph (the launches of a chunk) is replayed per chunk.           */
/* ------------------------------------------------------------------ */
typedef struct pt_batch pt_batch;
pt_status pt_batch_create(pt_batch **out);
void pt_batch_destroy(pt_batch *batch);
/* Adds a render of a committed scene (the scene must outlive the batch;
 * one scene may appear in several renders).  All renders of a batch use
 * one estimator.  PT_RENDER_KEEP_SAMPLES is not supported.                 */
pt_status pt_batch_add(pt_batch *batch, pt_scene *scene, const pt_render_cfg *cfg,
                       uint32_t *out_index);
/* Changes the configuration of render `index` (e.g. its seed).            */
pt_status pt_batch_set_cfg(pt_batch *batch, uint32_t index, const pt_render_cfg *cfg);

enum {
  PT_BATCH_NO_GRAPH = 1u << 8    /* launch eagerly instead of replaying a graph */
};
typedef struct pt_batch_cfg {
  uint32_t wavefront;   /* path slots in flight (0 = the batch tuning, else 64M) */
  uint32_t flags;       /* PT_RENDER_COUNTERS, PT_RENDER_SORT_HITS, PT_RENDER_TIMING (eager only),
                           PT_BATCH_NO_GRAPH                                    */
} pt_batch_cfg;
/* Renders every render of the batch into its own film.  stats sums over
 * the renders.                                                            */
pt_status pt_batch_render(pt_batch *batch, const pt_batch_cfg *cfg, pt_stats *stats);
/* Film of render `index` (width*height*3 per-pixel means, as pt_film_read). */
pt_status pt_batch_film_read(pt_batch *batch, uint32_t index, float *out_rgb);
pt_status pt_batch_set_tuning(pt_batch *batch, const pt_tuning *tuning);
/* Picks the wavefront size, grid sizes, sort-key bits and (PT_TUNE_LEAF)
 * one BVH leaf size for the batch as pt_autotune does for one render; the
 * leaf size re-commits the batch's scenes with it.                       */
pt_status pt_batch_autotune(pt_batch *batch, const pt_autotune_cfg *tc, pt_tuning *out);

/* ------------------------------------------------------------------ */
/* Learned path guiding (no counterpart in the reference: its           */
/* integrators sample the BSDF and the lights only,                     */
/* integrator/maxdepth_nee.hpp:9-256).  A small MLP maps a surface      */
/* point (position in the scene box, normal) to a mixture of 8 von      */
/* Mises-Fisher lobes over incident directions and the probability c of */
/* sampling it instead of the BSDF.  The network has BF16 weights and   */
/* FP32 accumulation, runs on the matrix cores, and is trained online   */
/* with Adam on the radiance the renderer's own paths carry.  Guided    */
/* vertices weight every direction by the one-sample MIS pdf            */
/* c p_guide + (1 - c) p_bsdf, and NEE uses that pdf in its power       */
/* heuristic, so the estimator stays unbiased for any network state.    */
/* ------------------------------------------------------------------ */
typedef struct pt_guide pt_guide;
#define PT_GUIDE_INPUTS 6      /* position in [0,1]^3, (normal + 1) / 2        */
#define PT_GUIDE_OUTPUTS 41    /* raw outputs (see pathtracer_amd/csrc/pt_guide.h) */
#define PT_GUIDE_NPARAMS 27497 /* FP32 parameters: per layer W[k][n] then b[n]   */
#define PT_GUIDE_MIX 48        /* floats of a processed mixture                */

/* A guide for a committed scene (its bounding box normalises positions),
 * with seeded initial weights.                                            */
pt_status pt_guide_create(pt_scene *scene, uint32_t seed, pt_guide **out);
void pt_guide_destroy(pt_guide *guide);
pt_status pt_guide_get_params(pt_guide *guide, float *params);
/* Sets the FP32 parameters (BF16 copies follow) and resets Adam.          */
pt_status pt_guide_set_params(pt_guide *guide, const float *params);
/* The network on n points x (n*6): raw outputs (n*41) and/or processed
 * mixtures (n*48: per lobe (kappa mu, log-weight), the 8 weights, c).      */
pt_status pt_guide_forward(pt_guide *guide, const float *x, size_t n, float *raw,
                           float *mix);
/* On n processed mixtures: guide pdf at dirs (n*3) and/or samples from
 * uniforms u (n*3) with their pdf.                                        */
pt_status pt_guide_eval(const float *mix, const float *dirs, const float *u, size_t n,
                        float *pdf, float *sample_dir, float *sample_pdf);
/* Training samples: n*11 floats (x[6], direction[3], p_bsdf, weight F/q).
 * Mean loss -w (log p_guide + log(c p_guide + (1 - c) p_bsdf)) and its gradient
 * (PT_GUIDE_NPARAMS floats) over the n samples.                           */
pt_status pt_guide_gradient(pt_guide *guide, const float *samples, size_t n, float *grad,
                            float *loss);
/* The range of the selection probability c = c_min + (c_max - c_min)
 * sigmoid(s) (default [0.05, 0.5]): 1 / (1 - c_max) bounds how much a guided
 * vertex can raise a path's throughput over BSDF sampling.                 */
pt_status pt_guide_set_select_range(pt_guide *guide, float c_min, float c_max);
/* The training samples the last training pass of pt_render_guided
 * collected (n*11 floats, the layout above): copies min(n, cap) of them.  */
pt_status pt_guide_samples(pt_guide *guide, float *out, size_t cap, size_t *n);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the n samples.            */
pt_status pt_guide_train_step(pt_guide *guide, const float *samples, size_t n, float lr,
                              float *loss);

enum {
  PT_GUIDE_NO_TRAIN = 1u << 0,   /* render with the current network only     */
  PT_GUIDE_PATH_QUERIES = 1u << 1 /* query the network at every path's camera
                                    vertex; by default the camera vertices of
                                    a pixel share one query (at the hit of
                                    the pixel's first sample in each chunk),
                                    which leaves the estimator unbiased     */
};
typedef struct pt_guide_cfg {
  uint32_t train_spp;      /* samples per pixel rendered while training
                              (0 = a sixteenth of the frame's)            */
  uint32_t first_pass_spp; /* training passes: f, 2f, 4f, ... spp (0: f = 4) */
  uint32_t batch;          /* training batch (0 = 16384)                    */
  uint32_t max_steps;      /* Adam steps after each training pass (0 = 96)  */
  uint32_t record_stride;  /* one path in this many records (0 = 1)         */
  uint32_t record_verts;   /* vertices recorded per path (0 = guide_verts)  */
  uint32_t kinds;          /* material kinds guided, bit per pt_mat_kind
                              (0 = Lambertian, rough conductor, Disney)     */
  float lr;                /* Adam learning rate (0 = 1e-3)                 */
  uint32_t flags;          /* PT_GUIDE_*                                    */
  uint32_t guide_verts;    /* vertices 0 .. guide_verts-1 are guided (0 = 1) */
} pt_guide_cfg;
typedef struct pt_guide_stats {
  uint32_t passes, steps;
  uint64_t samples;        /* training samples collected                    */
  float last_loss;
  double ms_train;         /* Adam steps (device time)                      */
} pt_guide_stats;
/* Renders cfg's frame (estimator PT_EST_NEE_MIS) with the guide: training
 * passes of growing spp, each followed by Adam steps on its samples, then
 * the rest of the frame with the trained network.  Every pass is unbiased,
 * so all count equally in the film.  PT_RENDER_KEEP_SAMPLES is supported
 * with PT_GUIDE_NO_TRAIN only (the frame is then one pass).                */
pt_status pt_render_guided(pt_scene *scene, const pt_render_cfg *cfg, pt_guide *guide,
                           const pt_guide_cfg *gcfg, pt_stats *stats,
                           pt_guide_stats *gstats);

/* ------------------------------------------------------------------ */
/* Parity / test entry points (batched, executed on the device).        */
/* ------------------------------------------------------------------ */

/* HaltonSampler::SampleDimension (sampler.hpp:54-73) and the other
 * samplers.  For Halton kinds: out[i] = value of dimension dim[i] at
 * sequence index index[i].  For Sobol/stratified: index[i] is the sample
 * number inside pixel[i], spp is the samples per pixel.                   */
pt_status pt_sample(uint32_t sampler, uint32_t seed, uint32_t spp,
                    const uint32_t *dim, const uint64_t *index,
                    const uint32_t *pixel, size_t n, float *out);

/* BVHNode::intersect (primitives/bvh.hpp:27-41) over the committed scene.
 * rays: n*8 floats (ox,oy,oz,dx,dy,dz,tmin,tmax).  prim: primitive id
 * (triangles first in insertion order, then spheres, then rects, then 8
 * segments per curve) or -1.                                               */
pt_status pt_intersect(pt_scene *scene, const float *rays, size_t n,
                       int any_hit, int32_t *prim, float *t);

/* Material::f and Material::scatter_pdf(wo,wi,rec) for one material over n
 * tuples.  wo, wi, nrm: n*3; is_inside: n bytes (HitRecord::is_inside).     */
pt_status pt_bsdf_eval(const pt_material_desc *mat, const float *wo,
                       const float *wi, const float *nrm,
                       const uint8_t *is_inside, size_t n, float *f,
                       float *pdf);
/* Material::scatter with explicit samples (uc, u2).  Outputs: wi (n*3),
 * attenuation (n*3), pdf (ScatterRecord::pdf_val), flags (BxDFFlags of the
 * sampled lobe), ok (scatter returned true).  aux: one extra uniform per
 * tuple for lobe choices the reference draws from rand_num (Phong).       */
pt_status pt_bsdf_sample(const pt_material_desc *mat, const float *wo,
                         const float *nrm, const uint8_t *is_inside,
                         const float *uc, const float *u2, const float *aux,
                         size_t n, float *wi, float *f, float *pdf,
                         uint32_t *flags, uint8_t *ok);

/* Light::get_sample / get_normal_at / pdf_value for light `light` of a
 * committed scene (light/light.hpp:17-306, light/envLight.hpp:20-76).
 * u2: n*2; ref: n*3 shading points.  Outputs p (n*3 point, or dir*1e5 for
 * the environment), nrm (n*3), pdf (solid-angle pdf of pdf_value(ref, p-ref)),
 * le (n*3, L_emit towards ref).                                             */
pt_status pt_light_sample(pt_scene *scene, uint32_t light, const float *u2,
                          const float *ref, size_t n, float *p, float *nrm,
                          float *pdf, float *le);

/* The environment's importance sampling as the NEE+MIS estimator uses it
 * (2-D piecewise-constant distribution; the reference samples its
 * environment uniformly, texture/texture.hpp:120-131): u2 n*2 -> unit
 * direction dir (n*3) and its solid-angle pdf (n).                          */
pt_status pt_env_sample(pt_scene *scene, const float *u2, size_t n, float *dir, float *pdf);

pt_status pt_device_synchronize(void);
const char *pt_last_error(void);
const char *pt_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PATHTRACER_AMD_H */
