// pt_guide_host.h -- the host side of a learned guide (pt_guide.cpp), shared
// with the guided render loop (pt_host.cpp).
#pragma once
#include "pt_common.h"
#include "pt_guide_launch.h"

struct pt_guide {
  int device = -1;
  hipStream_t stream = nullptr;  // the guide's own entry points
  float lo[3] = {0.f, 0.f, 0.f}, inv_ext[3] = {1.f, 1.f, 1.f};
  float c_lo = 0.05f, c_span = 0.45f;  // selection probability in [c_lo, c_lo + c_span]
  // parameters: FP32 master, Adam moments, the kernels' BF16 / padded copies
  DevBuf params, mom, vel, fwd, bwd, bias;
  uint32_t adam_t = 0;
  // training scratch (grown to the largest batch)
  DevBuf xt, dzt, gpart, dbpart, losspart, bsum;
  uint32_t ld = 0, splits = 8;
  // samples collected by the guided renderer, and their count
  DevBuf tset, tcount;
  uint32_t tcap = 0;
  // scratch of the test entry points
  DevBuf io_a, io_b, io_c;

  pt::GuideView view() const {
    pt::GuideView v{};
    v.fwd = fwd.as<uint16_t>();
    v.bwd = bwd.as<uint16_t>();
    v.bias = bias.as<float>();
    for (int a = 0; a < 3; ++a) {
      v.lo[a] = lo[a];
      v.inv_ext[a] = inv_ext[a];
    }
    v.c_lo = c_lo;
    v.c_span = c_span;
    return v;
  }
  pt::GuideParamBufs pbufs() const {
    pt::GuideParamBufs b{};
    b.params = params.as<float>();
    b.mom = mom.as<float>();
    b.vel = vel.as<float>();
    b.fwd = fwd.as<uint16_t>();
    b.bwd = bwd.as<uint16_t>();
    b.bias = bias.as<float>();
    return b;
  }
  pt::GuideTrainBufs tbufs() const {
    pt::GuideTrainBufs t{};
    t.ld = ld;
    t.splits = splits;
    t.xt = xt.as<uint16_t>();
    t.dzt = dzt.as<uint16_t>();
    t.gpart = gpart.as<float>();
    t.dbpart = dbpart.as<float>();
    t.losspart = losspart.as<float>();
    t.bsum = bsum.as<float>();
    return t;
  }
};

// Scene facts the guide needs (pt_host.cpp): bounding box and device.
pt_status ptx_scene_box(pt_scene *s, float lo[3], float hi[3], int *device);
// Grows the training scratch to batches of `batch` samples.
pt_status ptx_guide_reserve(pt_guide *g, uint32_t batch);
// One Adam step on `batch` samples picked from samples[0, n) (seed 0: the
// first `batch` in order), on stream st.  loss: mean loss (synchronises) or null.
pt_status ptx_guide_step(pt_guide *g, hipStream_t st, const float *samples, uint32_t n, uint32_t batch,
                         uint32_t pick_seed, float lr, float *loss);
