// Host-side pieces shared by the library's translation units: error types,
// the stream-ordered copy, the per-handle device arena, box validation and
// the stepper interface behind pf_handle.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pfhip.h"
#include "pfhip_internal.h"

namespace pfhost {

struct HipError {
  int code;
  std::string msg;
};

#define HIP_CHECK(expr)                                                                    \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      throw HipError{PF_E_HIP, std::string("HIP error: ") + hipGetErrorString(_e) + " (" + \
                                   #expr + ")"};                                           \
  } while (0)

struct ArgError {
  std::string msg;
};

// Host <-> device copy ordered on stream s (after the work already queued
// there, such as the clearing of a fresh allocation) that returns once it is
// complete, so the host buffer may be a temporary. The handles' streams are
// non-blocking, so a plain hipMemcpy on the null stream is not ordered
// against them.
inline void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  if (bytes == 0) return;
  HIP_CHECK(hipMemcpyAsync(dst, src, bytes, kind, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

// Device allocations of one handle, freed together.
class DeviceArena {
 public:
  ~DeviceArena() { release(); }
  void* alloc(size_t bytes) {
    if (bytes == 0) return nullptr;
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, bytes));
    ptrs_.push_back(p);
    return p;
  }
  template <typename T>
  T* upload(const T* host, size_t n, hipStream_t s) {
    if (n == 0) return nullptr;
    T* d = static_cast<T*>(alloc(n * sizeof(T)));
    copy_sync(d, host, n * sizeof(T), hipMemcpyHostToDevice, s);
    return d;
  }
  template <typename T>
  T* zeros(size_t n, hipStream_t s) {
    if (n == 0) return nullptr;
    T* d = static_cast<T*>(alloc(n * sizeof(T)));
    HIP_CHECK(hipMemsetAsync(d, 0, n * sizeof(T), s));
    return d;
  }
  void release() {
    for (void* p : ptrs_) (void)hipFree(p);
    ptrs_.clear();
  }
  // Free one allocation of this arena (no work may still use it).
  void free(void* p) {
    auto it = std::find(ptrs_.begin(), ptrs_.end(), p);
    if (it == ptrs_.end()) return;
    HIP_CHECK(hipFree(p));
    ptrs_.erase(it);
  }

 private:
  std::vector<void*> ptrs_;
};

inline pfhip::BoxDev check_box(const pf_box& b, const int64_t shape[3]) {
  pfhip::BoxDev out;
  for (int a = 0; a < 3; ++a) {
    if (b.count[a] < 1 || b.stride[a] < 1 || b.start[a] < 0 ||
        b.start[a] + (b.count[a] - 1) * b.stride[a] >= shape[a])
      throw ArgError{"box exceeds the grid on axis " + std::to_string(a)};
    out.start[a] = b.start[a];
    out.count[a] = b.count[a];
    out.stride[a] = b.stride[a];
  }
  return out;
}


}  // namespace pfhost

struct pf_handle {
  virtual ~pf_handle() = default;
  virtual void advance(int64_t step0, int64_t n) = 0;
  virtual void read_box(int comp, const pf_box& box, void* dst) = 0;
  virtual void read_box_q(int comp, const pf_box& box, int bits, void* dst, double* scale) = 0;
  virtual void write_field(int comp, const void* src) = 0;
  virtual void dft_read(int mon, int slot, double* dst) = 0;
  virtual void flux_read(double* dst) = 0;
  virtual void timing_enable(int every) = 0;   // 0 off, k: step groups of every k-th step
  virtual void checkpoint_enable(int interval) = 0;
  virtual void adjoint(int n_cot, const pf_cotangent* cot, double* grad) = 0;
  virtual void timing_read(double* ms, int64_t* launches) = 0;
  virtual void dft_flush() = 0;
  virtual void info(pf_info* out) const = 0;
  virtual void field_view(int comp, pf_view* out) = 0;
  virtual void dft_view(int mon, pf_view* out) = 0;
  int64_t shape[3];
};


namespace pfhost {
// The beyond-HBM stepper (stream_stepper.cpp) behind pf_create with
// PF_FLAG_STREAM.
pf_handle* make_streamer(const pf_desc& d);
}  // namespace pfhost
