// Host-side helpers shared by the engine's translation units: error handling
// behind the C ABI's status codes, device buffers, and the job checks used by
// both the one-shot and the served entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/totem_abi.h"

namespace totem {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Message of the last failure on the calling thread.
std::string& last_error();

#define HIP_OK(expr)                                                                                     \
  do {                                                                                                   \
    hipError_t e_ = (expr);                                                                              \
    if (e_ != hipSuccess)                                                                                \
      throw ::totem::Error(e_ == hipErrorOutOfMemory ? TOTEM_ERR_MEMORY : TOTEM_ERR_DEVICE,              \
                           std::string(#expr) + " failed: " + hipGetErrorString(e_));                   \
  } while (0)

inline void require(bool ok, const std::string& msg, int code = TOTEM_ERR_ARG) {
  if (!ok) throw Error(code, msg);
}

// Runs f, turning exceptions into a status code and the thread's last error.
template <class F> int guarded(F&& f) {
  try {
    f();
    last_error().clear();
    return TOTEM_OK;
  } catch (const Error& e) {
    last_error() = e.what();
    return e.code;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return TOTEM_ERR_DEVICE;
  }
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  void alloc(size_t b) {
    release();
    if (b == 0) return;
    HIP_OK(hipMalloc(&p, b));
    bytes = b;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  template <class T> T* as() const { return static_cast<T*>(p); }
};

inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Argument checks of a device STEM job (sizes, pointers, scan windows inside the
// potential unless wrap is set).
void validate_job(const totem_stem_job& j);
// Mid-layer walk of the reference (cuda_source_lobato.cu:1340-1385): mos[s] is the
// mid layer written after slice s (-1: none), hit[l] = 1 if layer l is written.
void mid_walk(const totem_stem_job& j, std::vector<int>& mos, std::vector<int>& hit);
// Probes in flight whose waves fit the memory-side cache (the default batch).
int preferred_batch(int n);
// Pass variant (plan) of n-point grids in use on the current device (0 unless set).
int current_plan(int n);
// The library's own stream on the current device (used when the caller passes none).
hipStream_t lib_stream();
// Batch HRTEM imaging, transfer-function imaging and weighted patterns (hrtem.hip).
void hrtem_images_impl(const totem_hrtem_desc& d, double* paras_out, double* images, hipStream_t st);
void tcc_image_impl(int N, int M, const double* wave, const double* tcc, const double* ap, const double* gfsf,
                    double* image, hipStream_t st);
void weighted_pattern_impl(int N, int M, const void* fields, const double* weights, int shift, double* out,
                           hipStream_t st);
// Window origin (row, column) of scan position d (periodic when j.wrap).
void window_origin(const totem_stem_job& j, int d, int& r0, int& c0);

}  // namespace totem
