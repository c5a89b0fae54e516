// Device helpers shared by the generic and batched per-node Yee kernels.
#pragma once

#include "pfhip_internal.h"

namespace pfhip {

// Flat psi index of a compact PML array whose axis `ax` holds only the
// `pc` nodes with a non-zero CPML c coefficient (rust/src/kernel.rs:92-103).
__device__ __forceinline__ int64_t psi_index(int ax, int m, int i, int j, int k, int ny,
                                             int nz, int pc) {
  if (ax == 0) return ((int64_t)m * ny + j) * nz + k;
  if (ax == 1) return ((int64_t)i * pc + m) * nz + k;
  return ((int64_t)i * ny + j) * pc + m;
}

template <typename real, bool EXACT>
__device__ __forceinline__ real diff(real hi, real lo, real d, real inv_d) {
  if constexpr (EXACT) return (hi - lo) / d;
  else return (hi - lo) * inv_d;
}

// One derivative term of a curl: present flag, value, and PML state.
template <typename real>
struct Term {
  bool on;
  real d;
  real* psi;    // nullptr when the node is outside the PML slab
  real* psw;    // where the updated psi is written (psi, or the out-of-place array)
  int64_t pid;
  real b, c;
};

// One curl component: f + K (dA - dB) with the PML slab terms after the bulk
// term (same order as the generic kernels).
template <typename real>
__device__ __forceinline__ real curl_update(real f, real K, bool onA, real dA, bool pmlA,
                                            real psiA, real bA, real cA, real* newA, bool onB,
                                            real dB, bool pmlB, real psiB, real bB, real cB,
                                            real* newB) {
  if (!onA && !onB) return f;
  real cb;
  if (onA) {
    cb = dA;
    if (onB) cb = cb - dB;
  } else {
    cb = -dB;
  }
  f = f + cb * K;
  if (onA && pmlA) {
    const real p = bA * psiA + cA * dA;
    *newA = p;
    f = f + K * p;
  }
  if (onB && pmlB) {
    const real p = bB * psiB + cB * dB;
    *newB = p;
    f = f - K * p;
  }
  return f;
}

// CPML membership of an axis for the march kernels: node n is in the slab
// set iff n < lo or n >= hi; m(n) is its compact psi index.
template <typename real>
struct Slabs {
  int lo, hi;
  __device__ __forceinline__ bool in(int n) const { return n < lo || n >= hi; }
  __device__ __forceinline__ int m(int n) const { return n < lo ? n : lo + (n - hi); }
};

// Workgroup barrier ordering only LDS: the release/acquire fences are scoped
// to the local address space, so outstanding global loads (the next plane's
// prefetch) stay in flight across it. __syncthreads() would wait for them.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Load real at a 32-bit byte offset from a wave-uniform base (SGPR-base
// addressing; every in-plane offset of the march kernel is < 4 GiB).
template <typename real>
__device__ __forceinline__ real ld(const real* base, uint32_t boff) {
  return *reinterpret_cast<const real*>(reinterpret_cast<const char*>(base) + boff);
}
template <typename real>
__device__ __forceinline__ void st(real* base, uint32_t boff, real v) {
  *reinterpret_cast<real*>(reinterpret_cast<char*>(base) + boff) = v;
}
// Streaming store (nontemporal hint): for data not read again before it
// leaves the cache anyway, so it does not push out lines that are.
template <typename real>
__device__ __forceinline__ void st_nt(real* base, uint32_t boff, real v) {
  __builtin_nontemporal_store(v, reinterpret_cast<real*>(reinterpret_cast<char*>(base) + boff));
}

// Pointer in the global address space: accesses through it compile to
// global_* instructions. Generic (flat) accesses count against both the
// vector-memory and the LDS counters, which makes the compiler wait for
// every outstanding load wherever their paths merge with others.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gptr(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// Buffer-resource access: wave-uniform base, 32-bit lane byte offset. A
// lane offset of BUF_NONE lies beyond the resource's range, so that lane's
// load returns 0 and its store is dropped, without a memory access. Masking
// lanes this way instead of branching keeps the count of outstanding memory
// operations the same on every path, so the compiler can wait for exactly
// the loads it needs and leave the next plane's prefetch in flight.
constexpr uint32_t BUF_NONE = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7FFFFFFF, 0x00020000);
}
template <typename real>
__device__ __forceinline__ real bld(const real* base, uint32_t boff);
template <>
__device__ __forceinline__ float bld<float>(const float* base, uint32_t boff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(buf_rsrc(base), boff, 0, 0));
}
template <>
__device__ __forceinline__ double bld<double>(const double* base, uint32_t boff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(buf_rsrc(base), boff, 0, 0));
}
template <typename real>
__device__ __forceinline__ void bst(real* base, uint32_t boff, real v);
template <>
__device__ __forceinline__ void bst<float>(float* base, uint32_t boff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), buf_rsrc(base), boff, 0, 0);
}
template <>
__device__ __forceinline__ void bst<double>(double* base, uint32_t boff, double v) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), buf_rsrc(base), boff, 0, 0);
}
// Buffer store with the nontemporal cache policy (streaming output).
constexpr int BUF_NT = 2;
template <typename real>
__device__ __forceinline__ void bst_nt(real* base, uint32_t boff, real v);
template <>
__device__ __forceinline__ void bst_nt<float>(float* base, uint32_t boff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), buf_rsrc(base), boff, 0,
                                        BUF_NT);
}
template <>
__device__ __forceinline__ void bst_nt<double>(double* base, uint32_t boff, double v) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), buf_rsrc(base), boff, 0,
                                        BUF_NT);
}

// Buffer access through a resource made once (buf_rsrc) with a wave-uniform
// byte offset in the scalar operand: the address is base + soff + voff, and
// only voff is range-checked, so a lane with voff = BUF_NONE still does no
// access. Callers keep soff below 2^31 (and voff of a taking-part lane below
// 2^31), so base + soff + voff never wraps.
template <typename real>
__device__ __forceinline__ real blds(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float blds<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double blds<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <int AUX = 0>
__device__ __forceinline__ void bsts(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ void bsts(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, double v) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, AUX);
}

// Value of lane l+1 / l-1 of the wavefront (DPP wavefront shifts; the lane
// with no source keeps its own value).
__device__ __forceinline__ float lane_next(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_prev(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_next(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
  const int nlo = __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false);
  const int nhi = __builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
  return __longlong_as_double(((long long)nhi << 32) | (unsigned)nlo);
}
__device__ __forceinline__ double lane_prev(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
  const int nlo = __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false);
  const int nhi = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
  return __longlong_as_double(((long long)nhi << 32) | (unsigned)nlo);
}

}  // namespace pfhip
