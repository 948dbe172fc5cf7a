// pt_launch.h -- host-callable kernel launchers (defined in pt_kernels.hip).
#pragma once
#include <algorithm>
#include "pt_scene.h"

namespace pt {

void launch_generate(hipStream_t st, const DevScene &s, const DevCamera &c, const RenderParams &rp, const Wave &w,
                     const float *offx, const float *offy, const BatchView *b);
void launch_extend(hipStream_t st, const DevScene &s, const Wave &w, int qi, uint32_t max_n, uint32_t grid_cap,
                   bool count, const BatchView *b);
void launch_shade(hipStream_t st, const DevScene &s, const RenderParams &rp, const Wave &w, int qi, uint32_t d,
                  uint32_t max_n, uint32_t grid_cap, const BatchView *b, uint32_t mat_kinds);
void launch_shadow(hipStream_t st, const DevScene &s, const Wave &w, uint32_t max_n, uint32_t grid_cap, bool count,
                   bool visibility, const BatchView *b);
void launch_film(hipStream_t st, const RenderParams &rp, const Wave &w, float *film, float *samples,
                 unsigned long long *stats, const BatchView *b, uint32_t max_pix);
void launch_next_chunk(hipStream_t st, const BatchView &b);
// FrameBuffer::setOffsets (framebuffer.cpp:131-145) on the device: off[k] and
// off[spp + k] of sample k, the values the host computes for pt_render.
void launch_pixel_offsets(hipStream_t st, uint32_t spp_x, uint32_t spp_y, uint32_t seed, float *off);
// film[i] *= scale for the n floats of a film (per-pixel sums to means).
void launch_film_scale(hipStream_t st, float *film, size_t n, float scale);
void launch_set_chunk(hipStream_t st, const BatchView &b, uint32_t c);
void launch_sample(hipStream_t st, const DevScene &s, uint32_t sampler, uint32_t seed, uint32_t sx, uint32_t sy,
                   const uint32_t *dim, const uint64_t *index, const uint32_t *pixel, size_t n, float *out);
void launch_intersect(hipStream_t st, const DevScene &s, const float *rays, size_t n, int any, int32_t *prim,
                      float *t);
void launch_bsdf_eval(hipStream_t st, const pt_material_desc &m, const float *wo, const float *wi, const float *nrm,
                      const uint8_t *inside, size_t n, float *f, float *pdf);
void launch_bsdf_sample(hipStream_t st, const pt_material_desc &m, const float *wo, const float *nrm,
                        const uint8_t *inside, const float *uc, const float *u2, const float *aux, size_t n, float *wi,
                        float *f, float *pdf, uint32_t *flags, uint8_t *ok);
void launch_light_sample(hipStream_t st, const DevScene &s, uint32_t light, const float *u2, const float *ref,
                         size_t n, float *p, float *nrm, float *pdf, float *le);
void launch_env_sample(hipStream_t st, const DevScene &s, const float *u2, size_t n, float *dir, float *pdf);

}  // namespace pt
