// Device-side quantisation of FieldMonitor snapshots for compressed storage.
//
// The reference compresses a snapshot on the host (storage.py:128-141):
// scale = max |x| (1 when that is 0 or not finite), then x / scale * 127
// rounded half to even into int8 (8 bits) or x / scale rounded to float16
// (16 bits), in the snapshot's own precision. Here the max-abs reduction and
// the quantisation run on the GPU right after the snapshot gather, so one or
// two bytes per value cross PCIe instead of four or eight; the host only runs
// the byte codec. Same arithmetic, same bits as the reference's host code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "pfhip_internal.h"

namespace pfhip {

// max with NaN propagation: once a NaN is seen the result stays NaN (as
// numpy's max does), so such a snapshot falls back to scale 1.
__device__ __forceinline__ double nanmax(double m, double v) { return (v > m || v != v) ? v : m; }

template <typename real>
__global__ __launch_bounds__(256) void k_absmax_partial(const real* __restrict__ x, int64_t n,
                                                        double* __restrict__ part) {
  __shared__ double s[256];
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    m = nanmax(m, fabs((double)x[i]));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] = nanmax(s[threadIdx.x], s[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// IEEE binary16 bits of d rounded to nearest, ties to even, directly from
// the double (a float input is widened exactly first), with overflow to
// infinity from 65520 on and gradual underflow.
__device__ __forceinline__ uint16_t half_bits(double d) {
  const uint64_t u = (uint64_t)__double_as_longlong(d);
  const uint16_t sign = (uint16_t)((u >> 48) & 0x8000u);
  const uint64_t a = u & 0x7fffffffffffffffull;
  if (a > 0x7ff0000000000000ull) return sign | 0x7e00u;          // NaN
  if (a >= 0x40effe0000000000ull) return sign | 0x7c00u;         // |d| >= 65520 -> inf
  if (a == 0) return sign;
  const int e = (int)(a >> 52) - 1023;
  if (e < -25) return sign;                                       // below half the smallest subnormal
  const uint64_t m = (a & ((1ull << 52) - 1)) | (1ull << 52);     // 53-bit significand
  // keep q = m >> shift, round on the dropped bits
  const int shift = e >= -14 ? 42 : 28 - e;                       // normal : subnormal (43..53)
  uint64_t q = m >> shift;
  const uint64_t rem = m & ((1ull << shift) - 1), halfway = 1ull << (shift - 1);
  if (rem > halfway || (rem == halfway && (q & 1))) ++q;
  if (e >= -14) return sign | (uint16_t)(((e + 15) << 10) + (q - 1024));  // a carry bumps the exponent
  return sign | (uint16_t)q;                                      // q == 1024 is the smallest normal
}

template <typename real>
__global__ __launch_bounds__(256) void k_quantize(const real* __restrict__ x, int64_t n,
                                                  const double* __restrict__ part, int n_part,
                                                  int bits, void* __restrict__ out,
                                                  double* __restrict__ scale_out) {
  __shared__ double sc;
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int p = 0; p < n_part; ++p) m = nanmax(m, part[p]);
    if (m == 0.0 || !isfinite(m)) m = 1.0;
    sc = m;
    if (blockIdx.x == 0) *scale_out = m;
  }
  __syncthreads();
  const real s = (real)sc;   // exact: a value of the snapshot, or 1
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const real q = x[i] / s;
    if (bits == 8) static_cast<int8_t*>(out)[i] = (int8_t)rint(q * real(127));
    else static_cast<uint16_t*>(out)[i] = half_bits((double)q);
  }
}

template <typename real>
void launch_quantize(const real* x, int64_t n, int bits, double* part, int max_part, void* out,
                     double* scale, hipStream_t s) {
  const int blocks = (int)std::min<int64_t>(max_part, std::max<int64_t>(1, (n + 255) / 256));
  k_absmax_partial<real><<<blocks, 256, 0, s>>>(x, n, part);
  k_quantize<real><<<blocks, 256, 0, s>>>(x, n, part, blocks, bits, out, scale);
}
template void launch_quantize<float>(const float*, int64_t, int, double*, int, void*, double*,
                                     hipStream_t);
template void launch_quantize<double>(const double*, int64_t, int, double*, int, void*, double*,
                                      hipStream_t);

}  // namespace pfhip
