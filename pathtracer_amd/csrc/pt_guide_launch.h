// pt_guide_launch.h -- host-callable launchers of the guide kernels
// (defined in pt_guide.hip) and the device buffers they work on.
#pragma once
#include "pt_guide.h"
#include "pt_scene.h"

namespace pt {

// FP32 master parameters, Adam moments and the BF16 / padded copies the
// kernels read.
struct GuideParamBufs {
  float *params, *mom, *vel;
  uint16_t *fwd, *bwd;
  float *bias;
};
// Scratch of one training step over a batch (columns padded to ld, a
// multiple of 64): transposed layer inputs and output gradients, weight
// gradient partials of `splits` batch slices, bias gradient and loss partials
// per 16-sample tile.
struct GuideTrainBufs {
  uint32_t ld, splits;
  uint16_t *xt, *dzt;
  float *gpart, *dbpart, *losspart;
  float *bsum;  // reduced bias gradients (5 x 64)
};
#define PT_G_SAMPLE 12

// per_pixel: one query per pixel of the chunk (chunk_first, sppc, n: the
// chunk's first path sample, samples per pixel and path samples), at the hit
// of the pixel's first sample in the chunk; else one per live entry.
void launch_guide_infer(hipStream_t st, const DevScene &s, const Wave &w, int qi, const GuideView &gv,
                        uint32_t kinds, uint32_t grid, bool per_pixel, uint64_t chunk_first, uint32_t sppc,
                        uint32_t n);
void launch_guide_points(hipStream_t st, const GuideView &gv, const float *xs, size_t n, float *raw, float *mix);
void launch_guide_mix(hipStream_t st, const float *mix, const float *dirs, const float *u, size_t n, float *pdf,
                      float *out_dir, float *out_pdf);
void launch_guide_train(hipStream_t st, const GuideView &gv, const float *samples, uint32_t n_samples,
                        uint32_t batch, uint32_t pick_seed, const GuideTrainBufs &tb);
void launch_guide_update(hipStream_t st, const GuideParamBufs &pb, const GuideTrainBufs *tb, float lr, float b1,
                         float b2, float eps, float c1, float c2, int mode);
void launch_guide_collect(hipStream_t st, const Wave &w, uint32_t n_rec, uint32_t verts, uint32_t stride,
                          uint32_t pass, float *samples, uint32_t cap, uint32_t *count);

}  // namespace pt
