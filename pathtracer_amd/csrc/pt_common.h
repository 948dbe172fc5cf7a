// pt_common.h -- host-side helpers shared by the C ABI sources: error
// reporting, device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/pathtracer_amd.h"

// Records msg as the calling thread's last error (pt_last_error) and returns st.
pt_status pt_fail(pt_status st, const std::string &msg);
static inline pt_status fail(pt_status st, const std::string &msg) { return pt_fail(st, msg); }

#define HIP_TRY(expr)                                                                          \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess)                                                                      \
      return fail(PT_ERR_DEVICE, std::string(#expr " failed: ") + hipGetErrorString(_e));      \
  } while (0)

// One device allocation, grown on demand and freed with its owner.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;  // owns a device allocation: never copied
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    if (n == 0) n = 256;
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  template <class T>
  T *as() const { return reinterpret_cast<T *>(p); }
};

// Page-locked host memory (hipHostMalloc), for staging large uploads.
struct PinnedBuf {
  void *p = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    hipError_t e = hipHostMalloc(&p, n == 0 ? 256 : n, hipHostMallocDefault);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  template <class T>
  T *as() const { return reinterpret_cast<T *>(p); }
};

static inline hipError_t upload(DevBuf &b, const void *src, size_t n) {
  hipError_t e = b.alloc(n);
  if (e != hipSuccess) return e;
  if (n) return hipMemcpy(b.p, src, n, hipMemcpyHostToDevice);
  return hipSuccess;
}
