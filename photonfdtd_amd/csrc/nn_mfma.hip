// Training and inference kernels of the spectral surrogate on the gfx950
// matrix cores.
//
// Every convolution pass is an implicit GEMM on v_mfma_f32_16x16x32_bf16
// (BF16 operands, FP32 accumulation). A workgroup of four wavefronts owns a
// BM x BN tile of the output matrix (BM 128, BN 32/64/128 by the output
// width) and walks K in steps of 32: each step stages a BM x 32 slice of A
// and a BN x 32 slice of B in LDS (double-buffered, one barrier per step;
// the next step's global loads are in flight during the current step's
// MFMAs), and each wavefront issues (BM/32) x (BN/32) MFMAs on its quarter.
// Operands are read from the NHWC activations as 16-byte vectors of eight
// channels (no im2col buffer): along K where eight consecutive k are eight
// consecutive channels of one tap, otherwise along the tile's rows, written
// to LDS transposed. Layers whose channel counts are not multiples of eight
// (the first layer, cin = 1) take an element-wise path.
//
//   forward      Y[m = pixel][n = cout]   = sum_k X~[m][k = (tap, cin)] W[n][k]
//   data grad    dX[m = pixel][n = cin]   = sum_k dZ~[m][k = (tap, cout)] W[cout][tap][n]
//   weight grad  dW[cout][kk]             = sum_r X~[r][kk] dZ[r][cout]
//                (computed as C[m = kk = (tap, cin)][n = cout], split over r)
//
// X~ / dZ~ are the gathered (im2col / transposed-im2col) operands;
// dZ = dY * relu'(Y) is formed during the gather. Activations are stored as
// BF16 (raw 16-bit, round to nearest even), gradients as FP32 and rounded to
// BF16 when they enter an MFMA. Weights are kept in FP32 by the optimizer with
// a BF16 copy for the MFMAs.
//
// Operand lane layout of v_mfma_f32_16x16x32_bf16: lane l holds
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j = 0..7, and
// receives D[4 (l / 16) + j][l % 16], j = 0..3.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <cstdint>
#include <cstdio>

#include "../../include/pfhip.h"

namespace pfnn {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;   // NaN
  u += 0x7fffu + ((u >> 16) & 1u);                        // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Eight BF16 values packed in four dwords (element j in the low half of
// dword j/2 for even j).
__device__ __forceinline__ uint16_t half_of(const u4& v, int j) {
  return (uint16_t)(v[j >> 1] >> (16 * (j & 1)));
}
__device__ __forceinline__ void set_half(u4& v, int j, uint16_t h) {
  const uint32_t sh = 16 * (j & 1);
  v[j >> 1] = (v[j >> 1] & ~(0xffffu << sh)) | ((uint32_t)h << sh);
}
__device__ __forceinline__ u4 ld16(const void* p) { return *reinterpret_cast<const u4*>(p); }

struct Geom {
  int n, h, w, cin, cout, k, s, pad, ho, wo, kdim;
};

__host__ __device__ inline Geom make_geom(const pf_nn_conv& c) {
  Geom g;
  g.n = c.n;
  g.h = c.h;
  g.w = c.w;
  g.cin = c.cin;
  g.cout = c.cout;
  g.k = c.ksize;
  g.s = c.stride;
  g.pad = c.ksize / 2;
  g.ho = (c.h + 2 * g.pad - c.ksize) / c.stride + 1;
  g.wo = (c.w + 2 * g.pad - c.ksize) / c.stride + 1;
  g.kdim = c.ksize * c.ksize * c.cin;
  return g;
}

// Input pixel read by output pixel r through tap `tap`, or -1 (padding).
__device__ __forceinline__ int64_t in_pixel(const Geom& g, int r, int tap) {
  const int ky = tap / g.k, kx = tap - ky * g.k;
  const int ox = r % g.wo, t = r / g.wo;
  const int oy = t % g.ho, b = t / g.ho;
  const int iy = oy * g.s + ky - g.pad, ix = ox * g.s + kx - g.pad;
  if (iy < 0 || iy >= g.h || ix < 0 || ix >= g.w) return -1;
  return (int64_t)(b * g.h + iy) * g.w + ix;
}
// Output pixel that reads input pixel m through tap `tap`, or -1.
__device__ __forceinline__ int64_t out_pixel(const Geom& g, int m, int tap) {
  const int ky = tap / g.k, kx = tap - ky * g.k;
  const int ix = m % g.w, t = m / g.w;
  const int iy = t % g.h, b = t / g.h;
  const int ny = iy + g.pad - ky, nx = ix + g.pad - kx;
  if (ny < 0 || nx < 0 || ny % g.s || nx % g.s) return -1;
  const int oy = ny / g.s, ox = nx / g.s;
  if (oy >= g.ho || ox >= g.wo) return -1;
  return (int64_t)(b * g.ho + oy) * g.wo + ox;
}

// im2col element: output pixel r of the convolution, k = (tap, cin).
__device__ __forceinline__ uint16_t im2col(const Geom& g, const uint16_t* x, int r, int kk) {
  const int64_t p = in_pixel(g, r, kk / g.cin);
  return p < 0 ? (uint16_t)0 : x[p * g.cin + kk % g.cin];
}

// dZ = dY * relu'(Y) at output pixel r, channel co (Y is BF16 or FP32).
__device__ __forceinline__ float dz(const Geom& g, const float* dy, const void* y, int relu,
                                    int out_f32, int64_t r, int co) {
  const size_t i = (size_t)r * g.cout + co;
  float v = dy[i];
  if (relu) {
    const float yv = out_f32 ? static_cast<const float*>(y)[i]
                             : bf2f(static_cast<const uint16_t*>(y)[i]);
    if (!(yv > 0.f)) v = 0.f;
  }
  return v;
}
// Raw operand vectors. Loads are issued unconditionally (a lane with
// nothing to read loads element 0) and the validity flag and the dZ
// conversion are applied by finish() when the vector is staged, so a load
// under a lane-dependent condition never makes the wavefront wait for it
// at the condition's end.
struct RawV {
  u4 v;
  bool ok;
};
__device__ __forceinline__ u4 finish(const RawV& q) { return q.ok ? q.v : u4{0u, 0u, 0u, 0u}; }

// dZ = dY relu'(Y) of eight consecutive channels co..co+7 of output pixel
// r (co and cout multiples of 8): dY (FP32) and Y (BF16 or FP32) bytes.
struct RawDz {
  u4 d0, d1, y0, y1;
  bool ok;
  int relu, out_f32;   // relu < 0: d0 holds the finished BF16 values
};
// dZ[r][co..co+7] from FP32 dY and the layer's output Y, or from `dzb`, the
// same values already rounded to BF16 by k_make_dz (when not null).
__device__ __forceinline__ RawDz dz8_fetch(const Geom& g, const float* dy, const void* y, int relu,
                                           int out_f32, int64_t r, int co, bool ok,
                                           const uint16_t* dzb = nullptr) {
  RawDz q;
  q.ok = ok;
  q.relu = relu;
  q.out_f32 = out_f32;
  const size_t i = ok ? (size_t)r * g.cout + co : 0;
  if (dzb) {
    q.d0 = ld16(dzb + i);
    q.d1 = q.y0 = q.y1 = u4{0u, 0u, 0u, 0u};
    q.relu = -1;
    return q;
  }
  q.d0 = ld16(dy + i);
  q.d1 = ld16(dy + i + 4);
  q.y0 = q.y1 = u4{0u, 0u, 0u, 0u};
  if (relu) {
    if (out_f32) {
      q.y0 = ld16(static_cast<const float*>(y) + i);
      q.y1 = ld16(static_cast<const float*>(y) + i + 4);
    } else {
      q.y0 = ld16(static_cast<const uint16_t*>(y) + i);
    }
  }
  return q;
}
// As BF16 (round to nearest even).
__device__ __forceinline__ u4 finish(const RawDz& q) {
  if (q.relu < 0) return q.ok ? q.d0 : u4{0u, 0u, 0u, 0u};
  float v[8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    v[j] = __uint_as_float(q.d0[j]);
    v[4 + j] = __uint_as_float(q.d1[j]);
  }
  if (q.relu) {
    if (q.out_f32) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (!(__uint_as_float(q.y0[j]) > 0.f)) v[j] = 0.f;
        if (!(__uint_as_float(q.y1[j]) > 0.f)) v[4 + j] = 0.f;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (!(bf2f(half_of(q.y0, j)) > 0.f)) v[j] = 0.f;
    }
  }
  u4 out;
#pragma unroll
  for (int j = 0; j < 4; ++j) out[j] = (uint32_t)f2bf(v[2 * j]) | ((uint32_t)f2bf(v[2 * j + 1]) << 16);
  return q.ok ? out : u4{0u, 0u, 0u, 0u};
}
__device__ __forceinline__ u4 dz8(const Geom& g, const float* dy, const void* y, int relu,
                                  int out_f32, int64_t r, int co) {
  return finish(dz8_fetch(g, dy, y, relu, out_f32, r, co, true));
}

constexpr int BK = 32;        // K per MFMA step
constexpr int LDK = BK + 8;   // LDS row length (BF16): padding spreads the rows over the banks

// Running (tap, channel) of k = tap C + c, advanced by BK.
struct TapChan {
  int tap, c;
  __device__ __forceinline__ void set(int k, int C) {
    tap = k / C;
    c = k - tap * C;
  }
  __device__ __forceinline__ void step(int C) {
    c += BK;
    while (c >= C) {
      c -= C;
      ++tap;
    }
  }
};
// Running output pixel (b, oy, ox) of pixel index r, advanced by `by`.
struct PixIt {
  int b, oy, ox;
  __device__ __forceinline__ void set(int r, int ho, int wo) {
    ox = r % wo;
    const int t = r / wo;
    oy = t % ho;
    b = t / ho;
  }
  __device__ __forceinline__ void step(int by, int ho, int wo) {
    ox += by;
    while (ox >= wo) {
      ox -= wo;
      if (++oy == ho) {
        oy = 0;
        ++b;
      }
    }
  }
};
// (ky, kx) of tap t of a 1x1 or 3x3 kernel.
__device__ __forceinline__ void tap_yx(const Geom& g, int t, int& ky, int& kx) {
  ky = g.k == 3 ? t / 3 : 0;
  kx = t - ky * g.k;
}

// Operand access of the three passes. Each operand is read as 16-byte
// vectors of eight BF16 values, along k (A[m][k..k+7], B[k..k+7][n]) when
// the operand's ROWVEC flag is false, along the tile's rows (A[m..m+7][k],
// B[k][n..n+7]) when it is true, entries outside [0, M) x [0, K) zero. A
// thread's vectors keep their rows over the K walk, so a per-thread iterator
// decomposes the row (pixel or tap and channel) once (x_init) and x_next
// returns the vector at the current K step and advances k by BK with
// additions only. Layers whose channel counts are not multiples of eight
// take element-wise paths (the same values, one at a time).
struct FwdOp {   // M = output pixels, N = cout, K = (tap, cin)
  static constexpr bool A_ROWVEC = false, B_ROWVEC = false;
  static constexpr int MAX_BM = 128;
  Geom g;
  const uint16_t* x;
  const uint16_t* w;
  const float* bias;
  void* y;
  int relu, out_f32;
  bool vec;   // cin a multiple of 8: a k-group of eight is eight channels of one tap
  struct AIt {
    int k, m, b, iy0, ix0;
    bool ok;
    TapChan tc;
  };
  __device__ __forceinline__ AIt a_init(int m, int k, int M) const {
    AIt it;
    it.k = k;
    it.m = m;
    it.ok = m < M;
    const int mm = it.ok ? m : 0;
    const int ox = mm % g.wo, t = mm / g.wo, oy = t % g.ho;
    it.b = t / g.ho;
    it.iy0 = oy * g.s - g.pad;
    it.ix0 = ox * g.s - g.pad;
    it.tc.set(k, g.cin);
    return it;
  }
  using ARaw = RawV;
  __device__ __forceinline__ RawV a_next(AIt& it, int K) const {
    RawV q;
    if (vec) {
      int ky, kx;
      tap_yx(g, it.tc.tap, ky, kx);
      const int iy = it.iy0 + ky, ix = it.ix0 + kx;
      q.ok = it.ok && it.k < K && iy >= 0 && iy < g.h && ix >= 0 && ix < g.w;
      q.v = ld16(x + (q.ok ? ((size_t)(it.b * g.h + iy) * g.w + ix) * g.cin + it.tc.c : 0));
      it.tc.step(g.cin);
    } else {
      q.ok = true;
      q.v = u4{0u, 0u, 0u, 0u};
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (it.ok && it.k + j < K) set_half(q.v, j, im2col(g, x, it.m, it.k + j));
    }
    it.k += BK;
    return q;
  }
  struct BIt {
    int k;
    const uint16_t* row;
    bool ok;
  };
  __device__ __forceinline__ BIt b_init(int n, int k, int N) const {
    return BIt{k, w + (size_t)(n < N ? n : 0) * g.kdim, n < N};
  }
  using BRaw = RawV;
  __device__ __forceinline__ RawV b_next(BIt& it, int K) const {
    RawV q;
    if (vec) {
      q.ok = it.ok && it.k < K;
      q.v = ld16(it.row + (q.ok ? it.k : 0));
    } else {
      q.ok = true;
      q.v = u4{0u, 0u, 0u, 0u};
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (it.ok && it.k + j < K) set_half(q.v, j, it.row[it.k + j]);
    }
    it.k += BK;
    return q;
  }
  __device__ __forceinline__ void store(int m, int nn, float acc) const {
    float v = acc + (bias ? bias[nn] : 0.f);
    if (relu && !(v > 0.f)) v = 0.f;
    const size_t i = (size_t)m * g.cout + nn;
    if (out_f32) static_cast<float*>(y)[i] = v;
    else static_cast<uint16_t*>(y)[i] = f2bf(v);
  }
};

// One parity class of the input pixels of a data-gradient pass: pixels
// (s iy2 + py, s ix2 + px), and the taps through which an output pixel reads
// them (for stride 1: every pixel, every tap), four bits per tap in `taps`.
// The other taps would only multiply zeros.
struct DgClass {
  int py, px, hc, wc, ntap;
  uint64_t taps;
};

struct DataGradOp {   // M = input pixels of the class, N = cin, K = (tap of the class, cout)
  static constexpr bool A_ROWVEC = false, B_ROWVEC = true;
  static constexpr int MAX_BM = 128;
  Geom g;
  DgClass c;
  const uint16_t* w;
  const float* dy;
  const void* y;
  int relu, out_f32;
  float* dx;
  bool vec;   // cin and cout multiples of 8
  const uint16_t* dzb;   // dZ as BF16 (k_make_dz), or null: from dy and y
  __device__ __forceinline__ int tap_of(int i) const { return (int)((c.taps >> (4 * i)) & 15u); }
  __device__ __forceinline__ int pixel(int m) const {
    const int ix2 = m % c.wc, t = m / c.wc;
    const int iy2 = t % c.hc, b = t / c.hc;
    return (b * g.h + g.s * iy2 + c.py) * g.w + g.s * ix2 + c.px;
  }
  // output pixel reading input pixel (b, iy, ix) through tap t, or -1
  __device__ __forceinline__ int64_t out_of(int b, int iy, int ix, int t) const {
    int ky, kx;
    tap_yx(g, t, ky, kx);
    const int ny = iy + g.pad - ky, nx = ix + g.pad - kx;
    if (ny < 0 || nx < 0) return -1;
    const int oy = g.s == 2 ? ny >> 1 : ny, ox = g.s == 2 ? nx >> 1 : nx;
    if ((g.s == 2 && ((ny | nx) & 1)) || oy >= g.ho || ox >= g.wo) return -1;
    return (int64_t)(b * g.ho + oy) * g.wo + ox;
  }
  struct AIt {
    int k, b, iy, ix;
    bool ok;
    TapChan tc;
  };
  __device__ __forceinline__ AIt a_init(int m, int k, int M) const {
    AIt it;
    it.k = k;
    it.ok = m < M;
    const int mm = it.ok ? m : 0;
    const int ix2 = mm % c.wc, t = mm / c.wc, iy2 = t % c.hc;
    it.b = t / c.hc;
    it.iy = g.s * iy2 + c.py;
    it.ix = g.s * ix2 + c.px;
    it.tc.set(k, g.cout);
    return it;
  }
  using ARaw = RawDz;
  __device__ __forceinline__ RawDz a_next(AIt& it, int K) const {
    RawDz q;
    if (vec) {
      const int64_t r = it.ok && it.k < K ? out_of(it.b, it.iy, it.ix, tap_of(it.tc.tap)) : -1;
      q = dz8_fetch(g, dy, y, relu, out_f32, r < 0 ? 0 : r, it.tc.c, r >= 0, dzb);
      it.tc.step(g.cout);
    } else {
      // element-wise: the finished values in d0/d1 as FP32, no mask
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        v[j] = 0.f;
        const int kk = it.k + j;
        if (!it.ok || kk >= K) continue;
        const int64_t r = out_of(it.b, it.iy, it.ix, tap_of(kk / g.cout));
        if (r >= 0) v[j] = dz(g, dy, y, relu, out_f32, r, kk % g.cout);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        q.d0[j] = __float_as_uint(v[j]);
        q.d1[j] = __float_as_uint(v[4 + j]);
      }
      q.y0 = q.y1 = u4{0u, 0u, 0u, 0u};
      q.ok = true;
      q.relu = 0;
      q.out_f32 = out_f32;
    }
    it.k += BK;
    return q;
  }
  // B[k = (tap, co)][n..n+7] = W[co][tap][n..n+7]
  struct BIt {
    int k, n;
    bool ok;
    TapChan tc;
  };
  __device__ __forceinline__ BIt b_init(int n, int k, int N) const {
    BIt it;
    it.k = k;
    it.n = n;
    it.ok = n < N;
    it.tc.set(k, g.cout);
    return it;
  }
  using BRaw = RawV;
  __device__ __forceinline__ RawV b_next(BIt& it, int K) const {
    RawV q;
    const bool live = it.ok && it.k < K;
    const uint16_t* src = w + (live ? ((size_t)it.tc.c * g.k * g.k + tap_of(it.tc.tap)) * g.cin : 0);
    if (vec) {
      q.ok = live;
      q.v = ld16(src + (live ? it.n : 0));
    } else {
      q.ok = true;
      q.v = u4{0u, 0u, 0u, 0u};
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (live && it.n + j < g.cin) set_half(q.v, j, src[it.n + j]);
    }
    it.k += BK;
    it.tc.step(g.cout);
    return q;
  }
  __device__ __forceinline__ void store(int m, int nn, float acc) const {
    dx[(size_t)pixel(m) * g.cin + nn] = acc;
  }
};

struct WeightGradOp {   // M = (tap, cin), N = cout, K = output pixels (split)
  static constexpr bool A_ROWVEC = true, B_ROWVEC = true;
  static constexpr int MAX_BM = 128;   // (64-row tiles measured slower on every layer)
  Geom g;
  const uint16_t* x;
  const float* dy;
  const void* y;
  int relu, out_f32;
  float* part;          // [split][cout][kdim]: the partial sum of each pixel split
  bool vec;             // cin and cout multiples of 8
  const uint16_t* dzb;  // dZ as BF16 (k_make_dz), or null: from dy and y
  // A[m..m+7][r] = X~[r][m..m+7]
  struct AIt {
    int k, m, ky, kx, ci;
    bool ok;
    PixIt p;
  };
  __device__ __forceinline__ AIt a_init(int m, int k, int M) const {
    AIt it;
    it.k = k;
    it.m = m;
    it.ok = m < M;
    const int tap = (it.ok ? m : 0) / g.cin;
    it.ci = (it.ok ? m : 0) - tap * g.cin;
    tap_yx(g, tap, it.ky, it.kx);
    it.p.set(k, g.ho, g.wo);
    return it;
  }
  using ARaw = RawV;
  __device__ __forceinline__ RawV a_next(AIt& it, int K) const {
    RawV q;
    if (vec) {
      const int iy = it.p.oy * g.s - g.pad + it.ky, ix = it.p.ox * g.s - g.pad + it.kx;
      q.ok = it.ok && it.k < K && iy >= 0 && iy < g.h && ix >= 0 && ix < g.w;
      q.v = ld16(x + (q.ok ? ((size_t)(it.p.b * g.h + iy) * g.w + ix) * g.cin + it.ci : 0));
    } else {
      q.ok = true;
      q.v = u4{0u, 0u, 0u, 0u};
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (it.ok && it.k < K && it.m + j < g.kdim) set_half(q.v, j, im2col(g, x, it.k, it.m + j));
    }
    it.k += BK;
    it.p.step(BK, g.ho, g.wo);
    return q;
  }
  // B[r][n..n+7] = dZ[r][n..n+7]
  struct BIt {
    int k, n;
    bool ok;
  };
  __device__ __forceinline__ BIt b_init(int n, int k, int N) const { return BIt{k, n, n < N}; }
  using BRaw = RawDz;
  __device__ __forceinline__ RawDz b_next(BIt& it, int K) const {
    RawDz q;
    if (vec) {
      const bool live = it.ok && it.k < K;
      q = dz8_fetch(g, dy, y, relu, out_f32, live ? it.k : 0, it.n, live, dzb);
    } else {
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        v[j] = it.ok && it.k < K && it.n + j < g.cout ? dz(g, dy, y, relu, out_f32, it.k, it.n + j) : 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        q.d0[j] = __float_as_uint(v[j]);
        q.d1[j] = __float_as_uint(v[4 + j]);
      }
      q.y0 = q.y1 = u4{0u, 0u, 0u, 0u};
      q.ok = true;
      q.relu = 0;
      q.out_f32 = out_f32;
    }
    it.k += BK;
    return q;
  }
  __device__ __forceinline__ void store(int m, int nn, float acc) const {
    part[((size_t)blockIdx.z * g.cout + nn) * g.kdim + m] = acc;
  }
};

// Layers with one input channel (the design map): direct convolutions.
// K is the 9 (or 1) taps, far too short to fill an MFMA; these passes are
// bound by the output and gradient bytes. Every product is of two BF16
// values (exact in FP32) and sums are FP32, as in the MFMA passes.
//
// Forward: one workgroup per output row (b, oy); thread (ox, channel group
// of eight) sums over the taps in tap order; the weights (cout x taps) are
// staged in LDS as FP32.
__global__ __launch_bounds__(256) void k_conv_c1_fwd(Geom g, const uint16_t* x, const uint16_t* w,
                                                     const float* bias, void* y, int relu,
                                                     int out_f32) {
  extern __shared__ float sw[];   // [cout][taps]
  const int taps = g.k * g.k, groups = g.cout / 8;
  for (int i = threadIdx.x; i < g.cout * taps; i += 256) sw[i] = bf2f(w[i]);
  __syncthreads();
  const int row = blockIdx.x, oy = row % g.ho, b = row / g.ho;
  for (int idx = threadIdx.x; idx < g.wo * groups; idx += 256) {
    const int ox = idx % g.wo, co0 = 8 * (idx / g.wo);
    float xv[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      xv[t] = 0.f;
      if (t < taps) {
        int ky, kx;
        tap_yx(g, t, ky, kx);
        const int iy = oy * g.s - g.pad + ky, ix = ox * g.s - g.pad + kx;
        if (iy >= 0 && iy < g.h && ix >= 0 && ix < g.w) xv[t] = bf2f(x[(size_t)(b * g.h + iy) * g.w + ix]);
      }
    }
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      float a = 0.f;
#pragma unroll
      for (int t = 0; t < 9; ++t)
        if (t < taps) a += xv[t] * sw[(co0 + j) * taps + t];
      a += bias ? bias[co0 + j] : 0.f;
      v[j] = relu && !(a > 0.f) ? 0.f : a;
    }
    const size_t o = ((size_t)row * g.wo + ox) * g.cout + co0;
    if (out_f32) {
      float* yo = static_cast<float*>(y) + o;
#pragma unroll
      for (int j = 0; j < 8; ++j) yo[j] = v[j];
    } else {
      u4 pk;
#pragma unroll
      for (int j = 0; j < 4; ++j) pk[j] = (uint32_t)f2bf(v[2 * j]) | ((uint32_t)f2bf(v[2 * j + 1]) << 16);
      *reinterpret_cast<u4*>(static_cast<uint16_t*>(y) + o) = pk;
    }
  }
}

// Input gradient (one input channel, cout a multiple of 8 up to 64): a
// thread per (input pixel, group of eight channels); for every tap that
// reaches an output pixel it loads that pixel's eight dZ values (two 16-byte
// loads and one of Y for the ReLU mask; the loads of all taps issued before
// any is used) and adds their products with W[co][tap] (tap-major,
// channels in order); the groups of a pixel are added in group order
// through LDS. dZ is rounded to BF16 as in the MFMA passes.
__global__ __launch_bounds__(256) void k_conv_c1_dgrad(Geom g, const uint16_t* w, const float* dy,
                                                       const void* y, int relu, int out_f32,
                                                       float* dx) {
  extern __shared__ float sw[];   // [taps][cout], then the group partials
  const int taps = g.k * g.k, groups = g.cout / 8, per = 256 / groups;
  for (int i = threadIdx.x; i < g.cout * taps; i += 256) sw[(i % taps) * g.cout + i / taps] = bf2f(w[i]);
  float* red = sw + g.cout * taps;
  __syncthreads();
  const int grp = threadIdx.x % groups, pl = threadIdx.x / groups;
  const int64_t m = (int64_t)blockIdx.x * per + pl;
  const int64_t P = (int64_t)g.n * g.h * g.w;
  float acc = 0.f;
  if (pl < per && m < P) {
    const int ix = (int)(m % g.w), t = (int)(m / g.w), iy = t % g.h, b = t / g.h;
    int64_t rr[9];
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) {
      rr[tp] = -1;
      if (tp >= taps) continue;
      int ky, kx;
      tap_yx(g, tp, ky, kx);
      const int ny = iy + g.pad - ky, nx = ix + g.pad - kx;
      if (ny < 0 || nx < 0 || ny % g.s || nx % g.s) continue;
      const int oy = ny / g.s, ox = nx / g.s;
      if (oy < g.ho && ox < g.wo) rr[tp] = (int64_t)(b * g.ho + oy) * g.wo + ox;
    }
    RawDz d[9];
#pragma unroll
    for (int tp = 0; tp < 9; ++tp)
      d[tp] = dz8_fetch(g, dy, y, relu, out_f32, rr[tp] >= 0 ? rr[tp] : 0, 8 * grp, rr[tp] >= 0);
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) {
      if (tp >= taps) continue;
      const u4 dv = finish(d[tp]);
      const float* wt = sw + tp * g.cout + 8 * grp;
#pragma unroll
      for (int j = 0; j < 8; ++j) acc += bf2f(half_of(dv, j)) * wt[j];
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (grp == 0 && pl < per && m < P) {
    float s = red[threadIdx.x];
    for (int q = 1; q < groups; ++q) s += red[threadIdx.x + q];
    dx[m] = s;
  }
}

// Weight gradient, one input channel, cout <= 64: workgroup z sums output
// pixels [z PIX, (z + 1) PIX) into part[z][cout][taps]. Thread t handles
// channels t % 32 (+ 32) over the pixels of stream t / 32 (every eighth
// pixel, walked by additions); the loads of four pixels are issued before
// any is used; the eight streams are added in stream order through LDS, so
// the partials are the same bits on every run.
constexpr int C1_PIX = 512;
__global__ __launch_bounds__(256) void k_conv_c1_wgrad(Geom g, const uint16_t* x, const float* dy,
                                                       const void* y, int relu, int out_f32,
                                                       float* part) {
  __shared__ float sums[8][64 * 9];
  const int t = threadIdx.x, c = t & 31, q = t >> 5;
  const int taps = g.k * g.k;
  const int64_t R = (int64_t)g.n * g.ho * g.wo;
  const int64_t r0 = (int64_t)blockIdx.x * C1_PIX, r1 = min(R, r0 + C1_PIX);
  float acc[2][9];
#pragma unroll
  for (int bb = 0; bb < 2; ++bb)
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) acc[bb][tp] = 0.f;
  PixIt p;
  p.set((int)min(r0 + q, R - 1), g.ho, g.wo);
  constexpr int U = 4;
  for (int64_t rb = r0 + q; rb < r1; rb += 8 * U) {
    // addresses of four pixels (index -1: padding or past the range)
    int64_t xi[U][9];
    int64_t ri[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t r = rb + 8 * u;
      ri[u] = r < r1 ? r : -1;
#pragma unroll
      for (int tp = 0; tp < 9; ++tp) {
        int ky, kx;
        tap_yx(g, tp, ky, kx);
        const int iy = p.oy * g.s - g.pad + ky, ix = p.ox * g.s - g.pad + kx;
        xi[u][tp] = ri[u] >= 0 && tp < taps && iy >= 0 && iy < g.h && ix >= 0 && ix < g.w
                        ? (int64_t)(p.b * g.h + iy) * g.w + ix
                        : -1;
      }
      p.step(8, g.ho, g.wo);
    }
    // loads, unconditional (index 0 where there is nothing to read), so
    // none waits for another; the masks are applied below
    uint16_t xr[U][9];
    float dv[U][2], yv[U][2];
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int tp = 0; tp < 9; ++tp) xr[u][tp] = x[xi[u][tp] >= 0 ? xi[u][tp] : 0];
#pragma unroll
      for (int bb = 0; bb < 2; ++bb) {
        const int co = c + 32 * bb;
        const bool ok = ri[u] >= 0 && co < g.cout;
        const size_t i = ok ? (size_t)ri[u] * g.cout + co : 0;
        dv[u][bb] = dy[i];
        yv[u][bb] = relu ? (out_f32 ? static_cast<const float*>(y)[i]
                                    : bf2f(static_cast<const uint16_t*>(y)[i]))
                         : 1.f;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int bb = 0; bb < 2; ++bb) {
        const bool ok = ri[u] >= 0 && c + 32 * bb < g.cout;
        const float d = ok ? bf2f(f2bf(yv[u][bb] > 0.f ? dv[u][bb] : 0.f)) : 0.f;
#pragma unroll
        for (int tp = 0; tp < 9; ++tp) acc[bb][tp] += (xi[u][tp] >= 0 ? bf2f(xr[u][tp]) : 0.f) * d;
      }
  }
#pragma unroll
  for (int bb = 0; bb < 2; ++bb)
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) sums[q][(c + 32 * bb) * 9 + tp] = acc[bb][tp];
  __syncthreads();
  for (int i = t; i < g.cout * taps; i += 256) {
    const int co = i / taps, tp = i % taps;
    float s = sums[0][co * 9 + tp];
    for (int k = 1; k < 8; ++k) s += sums[k][co * 9 + tp];
    part[((size_t)blockIdx.x * g.cout + co) * taps + tp] = s;
  }
}

// Weight gradient, one input channel, cout a multiple of 4 (at most 64):
// the workgroup sums output pixels [z PIX, (z + 1) PIX) into
// part[z][cout][taps] like k_conv_c1_wgrad, with LP = 2^lp_log2 lanes per
// pixel: lane q of a pixel's group holds channels 4q..4q+3 (one 16-byte dY
// load, one 8- or 16-byte Y load per pixel), so a wave issues one dY load per
// 64 / LP pixels instead of one per two. Thread t walks the pixels
// r0 + t / LP, + 256 / LP, ... in order (U at a time, loads first); the
// pixel slots are added in slot order through LDS, so the partials are the
// same bits on every run. Indices fit 32 bits (host check).
constexpr int C1W_U = 2;
__global__ __launch_bounds__(256) void k_conv_c1_wgrad4(Geom g, int lp_log2, const uint16_t* x,
                                                        const float* dy, const void* y, int relu,
                                                        int out_f32, float* part) {
  __shared__ float sums[256 * 4 * 9];
  const int t = threadIdx.x;
  const int LP = 1 << lp_log2, PS = 256 >> lp_log2;
  const int c0 = 4 * (t & (LP - 1)), ps = t >> lp_log2;
  const bool cok = c0 < g.cout;
  const int taps = g.k * g.k;
  const int R = g.n * g.ho * g.wo;
  const int r0 = (int)blockIdx.x * C1_PIX, r1 = min(R, r0 + C1_PIX);
  float acc[4][9];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) acc[j][tp] = 0.f;
  PixIt p;
  p.set(min(r0 + ps, R - 1), g.ho, g.wo);
  for (int rb = r0 + ps; rb < r1; rb += C1W_U * PS) {
    int xi[C1W_U][9], di[C1W_U];
    bool ok[C1W_U];
#pragma unroll
    for (int u = 0; u < C1W_U; ++u) {
      const int r = rb + u * PS;
      ok[u] = r < r1 && cok;
      di[u] = ok[u] ? r * g.cout + c0 : 0;
#pragma unroll
      for (int tp = 0; tp < 9; ++tp) {
        int ky, kx;
        tap_yx(g, tp, ky, kx);
        const int iy = p.oy * g.s - g.pad + ky, ix = p.ox * g.s - g.pad + kx;
        xi[u][tp] = ok[u] && tp < taps && iy >= 0 && iy < g.h && ix >= 0 && ix < g.w
                        ? (p.b * g.h + iy) * g.w + ix
                        : -1;
      }
      p.step(PS, g.ho, g.wo);
    }
    // loads, unconditional (index 0 where there is nothing to read)
    uint16_t xr[C1W_U][9];
    float4 dv[C1W_U], yv[C1W_U];
#pragma unroll
    for (int u = 0; u < C1W_U; ++u) {
#pragma unroll
      for (int tp = 0; tp < 9; ++tp) xr[u][tp] = x[xi[u][tp] >= 0 ? xi[u][tp] : 0];
      dv[u] = *reinterpret_cast<const float4*>(dy + di[u]);
      if (!relu) {
        yv[u] = make_float4(1.f, 1.f, 1.f, 1.f);
      } else if (out_f32) {
        yv[u] = *reinterpret_cast<const float4*>(static_cast<const float*>(y) + di[u]);
      } else {
        const uint2 h = *reinterpret_cast<const uint2*>(static_cast<const uint16_t*>(y) + di[u]);
        yv[u] = make_float4(bf2f((uint16_t)(h.x & 0xFFFF)), bf2f((uint16_t)(h.x >> 16)),
                            bf2f((uint16_t)(h.y & 0xFFFF)), bf2f((uint16_t)(h.y >> 16)));
      }
    }
#pragma unroll
    for (int u = 0; u < C1W_U; ++u) {
      const float dj[4] = {dv[u].x, dv[u].y, dv[u].z, dv[u].w};
      const float yj[4] = {yv[u].x, yv[u].y, yv[u].z, yv[u].w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float d = ok[u] ? bf2f(f2bf(yj[j] > 0.f ? dj[j] : 0.f)) : 0.f;
#pragma unroll
        for (int tp = 0; tp < 9; ++tp) acc[j][tp] += (xi[u][tp] >= 0 ? bf2f(xr[u][tp]) : 0.f) * d;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int tp = 0; tp < 9; ++tp) sums[(t * 4 + j) * 9 + tp] = acc[j][tp];
  __syncthreads();
  const int width = 4 * LP;   // channels per pixel slot
  for (int i = t; i < g.cout * taps; i += 256) {
    const int co = i / taps, tp = i % taps;
    float s = sums[co * 9 + tp];
    for (int k = 1; k < PS; ++k) s += sums[(k * width + co) * 9 + tp];
    part[((size_t)blockIdx.x * g.cout + co) * taps + tp] = s;
  }
}

// dw[i] += sum over the splits of part[z][i], z = 0, 1, ... in order: the
// weight gradient is the same bits on every run (no atomics).
__global__ void k_reduce_splits(const float* part, int splits, int64_t count, float* dw) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = part[i];
  for (int z = 1; z < splits; ++z) s += part[(size_t)z * count + i];
  dw[i] += s;
}
// The same sum for many splits of a short gradient: one workgroup per
// element; thread t adds splits t, t + 256, ... in order, then a fixed
// tree over the threads (the same bits on every run).
__global__ __launch_bounds__(256) void k_reduce_splits_tree(const float* part, int splits,
                                                            int64_t count, float* dw) {
  __shared__ float red[256];
  const int64_t i = blockIdx.x;
  float s = 0.f;
  for (int z = threadIdx.x; z < splits; z += 256) s += part[(size_t)z * count + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) dw[i] += red[0];
}
inline void reduce_splits(const float* part, int splits, int64_t count, float* dw, hipStream_t s) {
  if (splits > 256 && count <= 65536)
    k_reduce_splits_tree<<<(unsigned)count, 256, 0, s>>>(part, splits, count, dw);
  else
    k_reduce_splits<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(part, splits, count, dw);
}

// C[M x N] (+)= A[M x K] B[K x N] over K range [z * kper, (z + 1) * kper).
template <class Op, int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm(Op op, int M, int N, int K, int kper) {
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][BM][LDK];
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][BN][LDK];
  constexpr int NVA = BM * BK / 8, NVB = BN * BK / 8;   // 16-byte vectors per slice
  constexpr int PA = (NVA + 255) / 256, PB = (NVB + 255) / 256;
  constexpr int WM = BM / 2, WN = BN / 2;                // each wave: a WM x WN quarter
  constexpr int FM = WM / 16, FN = WN / 16;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kb = blockIdx.z * kper, ke = min(K, kb + kper);
  const int t = threadIdx.x, l = t & 63, wv = t >> 6;
  const int wm = (wv >> 1) * WM, wn = (wv & 1) * WN;
  f4 acc[FM][FN];
  for (int i = 0; i < FM; ++i)
    for (int j = 0; j < FN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  typename Op::ARaw ra[PA];
  typename Op::BRaw rb[PB];
  // vector v of a slice: along k, (row v / 4, k-group v % 4); along the
  // rows, (k v % 32, row group v / 32), so a wave's LDS writes of one row
  // group are consecutive in k
  typename Op::AIt ia[PA];
  typename Op::BIt ib[PB];
#pragma unroll
  for (int p = 0; p < PA; ++p) {
    const int v = min(t + 256 * p, NVA - 1);
    ia[p] = Op::A_ROWVEC ? op.a_init(m0 + 8 * (v / BK), kb + v % BK, M)
                         : op.a_init(m0 + v / (BK / 8), kb + 8 * (v % (BK / 8)), M);
  }
#pragma unroll
  for (int p = 0; p < PB; ++p) {
    const int v = min(t + 256 * p, NVB - 1);
    ib[p] = Op::B_ROWVEC ? op.b_init(n0 + 8 * (v / BK), kb + v % BK, N)
                         : op.b_init(n0 + v / (BK / 8), kb + 8 * (v % (BK / 8)), N);
  }
  // the next K step's vectors (each call advances the iterators by BK)
  auto load = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int p = 0; p < PA; ++p) ra[p] = op.a_next(ia[p], ke);
#pragma unroll
    for (int p = 0; p < PB; ++p) rb[p] = op.b_next(ib[p], ke);
  };
  auto stage = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int p = 0; p < PA; ++p) {
      const int v = t + 256 * p;
      if (NVA % 256 != 0 && v >= NVA) continue;
      const u4 val = finish(ra[p]);
      if (Op::A_ROWVEC) {
#pragma unroll
        for (int j = 0; j < 8; ++j) sA[buf][8 * (v / BK) + j][v % BK] = half_of(val, j);
      } else {
        *reinterpret_cast<u4*>(&sA[buf][v / (BK / 8)][8 * (v % (BK / 8))]) = val;
      }
    }
#pragma unroll
    for (int p = 0; p < PB; ++p) {
      const int v = t + 256 * p;
      if (NVB % 256 != 0 && v >= NVB) continue;
      const u4 val = finish(rb[p]);
      if (Op::B_ROWVEC) {
#pragma unroll
        for (int j = 0; j < 8; ++j) sB[buf][8 * (v / BK) + j][v % BK] = half_of(val, j);
      } else {
        *reinterpret_cast<u4*>(&sB[buf][v / (BK / 8)][8 * (v % (BK / 8))]) = val;
      }
    }
  };
  const int nk = ke > kb ? (ke - kb + BK - 1) / BK : 0;
  if (nk > 0) {
    load();
    stage(0);
  }
  __syncthreads();
  for (int it = 0; it < nk; ++it) {
    const int buf = it & 1;
    if (it + 1 < nk) load();
    bf8 fa[FM], fb[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
      fa[i] = __builtin_bit_cast(bf8, *reinterpret_cast<const u4*>(&sA[buf][wm + 16 * i + (l & 15)][8 * (l >> 4)]));
#pragma unroll
    for (int j = 0; j < FN; ++j)
      fb[j] = __builtin_bit_cast(bf8, *reinterpret_cast<const u4*>(&sB[buf][wn + 16 * j + (l & 15)][8 * (l >> 4)]));
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    if (it + 1 < nk) stage(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (l >> 4) + r;
        const int nn = n0 + wn + 16 * j + (l & 15);
        if (m < M && nn < N) op.store(m, nn, acc[i][j][r]);
      }
}

// dZ = relu mask of dY, rounded to BF16 (the values every gradient pass
// multiplies), written once per layer so both GEMM passes read 2 bytes per
// value instead of 4 of dY plus Y. Thread t: values 8t..8t+7 (cout % 8 == 0).
__global__ __launch_bounds__(256) void k_make_dz(Geom g, const float* dy, const void* y, int relu,
                                                 int out_f32, uint16_t* dzb) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t R = (int64_t)g.n * g.ho * g.wo;
  const int per = g.cout / 8;
  if (v >= R * per) return;
  const int64_t r = v / per;
  const int co = (int)(v - r * per) * 8;
  *reinterpret_cast<u4*>(dzb + r * g.cout + co) = finish(dz8_fetch(g, dy, y, relu, out_f32, r, co, true));
}

// dbias[co] += sum over output pixels of dZ: one workgroup per channel.
__global__ __launch_bounds__(256) void k_bias_grad(Geom g, const float* dy, const void* y, int relu,
                                                   int out_f32, float* dbias) {
  __shared__ float part[256];
  const int co = blockIdx.x;
  const int R = g.n * g.ho * g.wo;
  float s = 0.f;
  for (int r = threadIdx.x; r < R; r += 256) s += dz(g, dy, y, relu, out_f32, r, co);
  part[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) dbias[co] += part[0];
}

// Mean squared error over n x d: loss[0] = mean (p - t)^2, dp = 2 (p - t) / (n d).
__global__ void k_mse(int total, const float* p, const float* tgt, float* dp, float* loss) {
  __shared__ float part[256];
  float s = 0.f;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256) {
    const float e = p[i] - tgt[i];
    dp[i] = 2.f * e / (float)total;
    s += e * e;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0] / (float)total);
}

// Adam with bias correction on FP32 master weights; refreshes the BF16 copy
// and clears the gradient.
__global__ void k_adam(int64_t n, float* p, float* g, float* m, float* v, uint16_t* pb, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (pb) pb[i] = f2bf(pi);
  g[i] = 0.f;
}

__global__ void k_to_bf16(int64_t n, const float* x, uint16_t* y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// Number of K ranges a GEMM of K = kk splits into: ceil(kk / kper) for
// kper = ceil(kk / splits) rounded up to BK.
inline int split_count(int kk, int splits) {
  const int kper = ((kk + splits - 1) / splits + BK - 1) / BK * BK;
  return (kk + kper - 1) / kper;
}

// Tile shape of a GEMM: BN (32, 64 or 128) by the output width N, BM (128,
// 64 or 32) the largest that still gives z K-ranges times the tile count
// 256 workgroups (one per CU), so small layers spread over the chip.
inline int tile_n(int N) { return N <= 32 ? 32 : N <= 64 ? 64 : 128; }
inline int tile_m(int M, int N, int z) {
  const int64_t nt = (N + tile_n(N) - 1) / tile_n(N);
  for (int bm : {128, 64})
    if ((int64_t)((M + bm - 1) / bm) * nt * z >= 256) return bm;
  return 32;
}

template <class Op, int BM>
void launch_gemm_bm(const Op& op, int M, int N, int K, int kper, int z, hipStream_t s) {
  const int bn = tile_n(N);
  dim3 grid((M + BM - 1) / BM, (N + bn - 1) / bn, z);
  if (bn == 32) k_gemm<Op, BM, 32><<<grid, 256, 0, s>>>(op, M, N, K, kper);
  else if (bn == 64) k_gemm<Op, BM, 64><<<grid, 256, 0, s>>>(op, M, N, K, kper);
  else k_gemm<Op, BM, 128><<<grid, 256, 0, s>>>(op, M, N, K, kper);
}

template <class Op>
void launch_gemm(const Op& op, int M, int N, int K, int splits, hipStream_t s) {
  const int kper = std::max(BK, ((K + splits - 1) / splits + BK - 1) / BK * BK);
  const int z = std::max(1, (K + kper - 1) / kper);   // K = 0: one range, the tile stores zeros
  const int bm = std::min(Op::MAX_BM, tile_m(M, N, z));
  if (bm == 128 && Op::MAX_BM >= 128) launch_gemm_bm<Op, Op::MAX_BM>(op, M, N, K, kper, z, s);
  else if (bm >= 64) launch_gemm_bm<Op, 64>(op, M, N, K, kper, z, s);
  else launch_gemm_bm<Op, 32>(op, M, N, K, kper, z, s);
}

// Largest weight-gradient workspace a conv backward may need from the
// library's static scratch array, in floats (16 MiB); the pixel splits of a
// layer are capped so that splits * cout * kdim stays within it.
constexpr int64_t SCRATCH_FLOATS = int64_t(4) << 20;

// Pixel splits of the weight-gradient GEMM of a layer (M = taps x cin,
// N = cout, K = output pixels): enough that the GEMM's tiles times the
// splits reach 512 workgroups, at least one per 8192 pixels, at most one
// per 256 pixels and at most SCRATCH_FLOATS / (cout kdim) (at least 1). A
// function of the shape only, so both backward entry points sum the same
// partials in the same order.
inline bool c1_wgrad(const Geom& g) { return g.cin == 1 && g.cout <= 64 && g.k <= 3; }
// the one-input-channel data gradient (k_conv_c1_dgrad)
inline bool c1_dgrad(const Geom& g) { return g.cin == 1 && g.cout % 8 == 0 && g.cout <= 64; }
// the four-channel form: 16-byte dY rows and 32-bit indices
inline bool c1_wgrad4(const Geom& g) {
  return c1_wgrad(g) && g.cout % 4 == 0 && (int64_t)g.n * g.ho * g.wo * g.cout < INT32_MAX &&
         (int64_t)g.n * g.h * g.w < INT32_MAX;
}
// Weight-gradient tiles of 288 rows (3x3 taps x 32 channels) where kdim is a
// multiple of 288 and cout <= 64: no padded rows, and dZ read once per 288
// rows of kdim instead of once per 128.
constexpr int WG_BM288 = 288;
inline bool wgrad_bm288(const Geom& g) {
  return g.kdim % WG_BM288 == 0 && g.cout <= 64 && g.cin % 8 == 0 && g.cout % 8 == 0;
}
inline int weight_grad_splits(const Geom& g) {
  const int64_t R = (int64_t)g.n * g.ho * g.wo;
  if (c1_wgrad(g)) return (int)((R + C1_PIX - 1) / C1_PIX);
  const int64_t per = (int64_t)g.cout * g.kdim;
  const int bm = wgrad_bm288(g) ? WG_BM288 : 128;
  const int64_t tiles = (int64_t)((g.kdim + bm - 1) / bm) * ((g.cout + tile_n(g.cout) - 1) / tile_n(g.cout));
  int64_t s = std::max<int64_t>((512 + tiles - 1) / tiles, (R + 8191) / 8192);
  s = std::min<int64_t>(s, std::max<int64_t>(1, R / 256));
  s = std::min<int64_t>(s, std::max<int64_t>(1, SCRATCH_FLOATS / per));
  return split_count((int)R, (int)s);
}

// Parity classes of a data-gradient pass (one for stride 1).
inline int dg_classes(const Geom& g, DgClass* out) {
  int n = 0;
  for (int py = 0; py < g.s; ++py)
    for (int px = 0; px < g.s; ++px) {
      DgClass c{};
      c.py = py;
      c.px = px;
      c.hc = (g.h - py + g.s - 1) / g.s;
      c.wc = (g.w - px + g.s - 1) / g.s;
      if (c.hc <= 0 || c.wc <= 0) continue;
      for (int ky = 0; ky < g.k; ++ky)
        for (int kx = 0; kx < g.k; ++kx)
          if ((py + g.pad - ky) % g.s == 0 && (px + g.pad - kx) % g.s == 0)
            c.taps |= (uint64_t)(ky * g.k + kx) << (4 * c.ntap++);
      out[n++] = c;
    }
  return n;
}

}  // namespace pfnn

namespace {

thread_local char nn_err[256];

int nn_fail(const char* what) {
  snprintf(nn_err, sizeof(nn_err), "%s", what);
  return PF_E_ARG;
}

int nn_status(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return PF_OK;
  snprintf(nn_err, sizeof(nn_err), "HIP error: %s (%s)", hipGetErrorString(e), what);
  return PF_E_HIP;
}

bool valid(const pf_nn_conv* c) {
  return c && c->n > 0 && c->h > 0 && c->w > 0 && c->cin > 0 && c->cout > 0 &&
         (c->ksize == 1 || c->ksize == 3) && (c->stride == 1 || c->stride == 2) &&
         (int64_t)c->n * c->h * c->w * (c->cin > c->cout ? c->cin : c->cout) < (int64_t(1) << 31);
}

}  // namespace

extern "C" {

const char* pf_nn_last_error(void) { return nn_err; }

int pf_nn_conv_forward(const pf_nn_conv* c, const void* x, const void* w, const float* bias,
                       void* y, void* stream) {
  if (!valid(c) || !x || !w || !y) return nn_fail("pf_nn_conv_forward: bad arguments");
  const pfnn::Geom g = pfnn::make_geom(*c);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g.cin == 1 && g.cout % 8 == 0 && g.cout <= 1024) {
    pfnn::k_conv_c1_fwd<<<(unsigned)(g.n * g.ho), 256, g.cout * g.k * g.k * sizeof(float), s>>>(
        g, static_cast<const uint16_t*>(x), static_cast<const uint16_t*>(w), bias, y, c->relu,
        c->out_f32);
    return nn_status("pf_nn_conv_forward");
  }
  pfnn::FwdOp op{g, static_cast<const uint16_t*>(x), static_cast<const uint16_t*>(w), bias, y,
                 c->relu, c->out_f32, g.cin % 8 == 0};
  pfnn::launch_gemm(op, g.n * g.ho * g.wo, g.cout, g.kdim, 1, s);
  return nn_status("pf_nn_conv_forward");
}

}  // extern "C"

namespace {

// Whether the layer's gradient GEMMs read dZ as BF16 from a prepass
// (vector operands on both channel counts, one of them a GEMM pass).
bool gemm_dz(const pfnn::Geom& g, bool dx, bool dw) {
  const bool vec = g.cin % 8 == 0 && g.cout % 8 == 0;
  const bool dx_gemm = dx && !pfnn::c1_dgrad(g);
  const bool dw_gemm = dw && !pfnn::c1_wgrad(g);
  return vec && (dx_gemm || dw_gemm);
}
size_t dz_bytes(const pfnn::Geom& g) {
  return ((size_t)g.n * g.ho * g.wo * g.cout * sizeof(uint16_t) + 255) / 256 * 256;
}

// The backward kernels of one layer; `part` is the weight-gradient workspace
// of weight_grad_splits(g) * cout * kdim floats (unused when dw is NULL);
// `dzb`, when not null, dz_bytes(g) of workspace for the BF16 dZ prepass
// (gemm_dz(g)): the same values the passes would otherwise round from dY
// and Y, so either way gives the same bits.
void conv_backward(const pf_nn_conv* c, const void* x, const void* w, const void* y,
                   const float* dy, float* dx, float* dw, float* dbias, float* part,
                   uint16_t* dzb, hipStream_t s) {
  const pfnn::Geom g = pfnn::make_geom(*c);
  if (dzb && gemm_dz(g, dx != nullptr, dw != nullptr)) {
    const int64_t nv = (int64_t)g.n * g.ho * g.wo * (g.cout / 8);
    pfnn::k_make_dz<<<(unsigned)((nv + 255) / 256), 256, 0, s>>>(g, dy, y, c->relu, c->out_f32, dzb);
  } else {
    dzb = nullptr;
  }
  if (dx && pfnn::c1_dgrad(g)) {
    const int64_t px = (int64_t)g.n * g.h * g.w, per = 256 / (g.cout / 8);
    pfnn::k_conv_c1_dgrad<<<(unsigned)((px + per - 1) / per), 256,
                            (g.cout * g.k * g.k + 256) * sizeof(float), s>>>(
        g, static_cast<const uint16_t*>(w), dy, y, c->relu, c->out_f32, dx);
  } else if (dx) {
    // one GEMM per parity class of the input pixels, over that class's taps
    pfnn::DgClass cls[4];
    const int ncls = pfnn::dg_classes(g, cls);
    for (int q = 0; q < ncls; ++q) {
      pfnn::DataGradOp op{g, cls[q], static_cast<const uint16_t*>(w), dy, y, c->relu, c->out_f32,
                          dx, g.cin % 8 == 0 && g.cout % 8 == 0, dzb};
      pfnn::launch_gemm(op, g.n * cls[q].hc * cls[q].wc, g.cin, cls[q].ntap * g.cout, 1, s);
    }
  }
  if (dw && pfnn::c1_wgrad(g)) {
    const int splits = pfnn::weight_grad_splits(g);
    const bool aligned = reinterpret_cast<uintptr_t>(dy) % 16 == 0 &&
                         reinterpret_cast<uintptr_t>(y) % (c->out_f32 ? 16 : 8) == 0;
    if (pfnn::c1_wgrad4(g) && aligned) {
      int lp_log2 = 0;
      while ((1 << lp_log2) * 4 < g.cout) ++lp_log2;
      pfnn::k_conv_c1_wgrad4<<<splits, 256, 0, s>>>(g, lp_log2, static_cast<const uint16_t*>(x), dy,
                                                     y, c->relu, c->out_f32, part);
    } else {
      pfnn::k_conv_c1_wgrad<<<splits, 256, 0, s>>>(g, static_cast<const uint16_t*>(x), dy, y,
                                                     c->relu, c->out_f32, part);
    }
    pfnn::reduce_splits(part, splits, (int64_t)g.cout * g.kdim, dw, s);
  } else if (dw) {
    const int splits = pfnn::weight_grad_splits(g);
    pfnn::WeightGradOp op{g, static_cast<const uint16_t*>(x), dy, y, c->relu, c->out_f32, part,
                          g.cin % 8 == 0 && g.cout % 8 == 0, dzb};
    const int K = g.n * g.ho * g.wo;
    if (pfnn::wgrad_bm288(g)) {
      const int kper = std::max(pfnn::BK, ((K + splits - 1) / splits + pfnn::BK - 1) / pfnn::BK * pfnn::BK);
      const int z = std::max(1, (K + kper - 1) / kper);
      pfnn::launch_gemm_bm<pfnn::WeightGradOp, pfnn::WG_BM288>(op, g.kdim, g.cout, K, kper, z, s);
    } else {
      pfnn::launch_gemm(op, g.kdim, g.cout, K, splits, s);
    }
    pfnn::reduce_splits(part, splits, (int64_t)g.cout * g.kdim, dw, s);
  }
  if (dbias) pfnn::k_bias_grad<<<g.cout, 256, 0, s>>>(g, dy, y, c->relu, c->out_f32, dbias);
}

}  // namespace

extern "C" {

int pf_nn_conv_backward(const pf_nn_conv* c, const void* x, const void* w, const void* y,
                        const float* dy, float* dx, float* dw, float* dbias, void* stream) {
  if (!valid(c) || !x || !w || !y || !dy) return nn_fail("pf_nn_conv_backward: bad arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // stream-ordered workspace (allocated, used and released in order on s):
  // the BF16 dZ, then the weight gradient's split partials
  const pfnn::Geom g = pfnn::make_geom(*c);
  const size_t zb = gemm_dz(g, dx != nullptr, dw != nullptr) ? dz_bytes(g) : 0;
  const size_t pb = dw ? (size_t)pfnn::weight_grad_splits(g) * g.cout * g.kdim * sizeof(float) : 0;
  char* ws = nullptr;
  if (zb + pb > 0 && hipMallocAsync(reinterpret_cast<void**>(&ws), zb + pb, s) != hipSuccess)
    return nn_status("pf_nn_conv_backward: workspace allocation");
  uint16_t* dzb = zb ? reinterpret_cast<uint16_t*>(ws) : nullptr;
  float* part = pb ? reinterpret_cast<float*>(ws + zb) : nullptr;
  conv_backward(c, x, w, y, dy, dx, dw, dbias, part, dzb, s);
  const int st = nn_status("pf_nn_conv_backward");
  if (ws) (void)hipFreeAsync(ws, s);
  return st;
}

}  // extern "C"

// Weight-gradient workspace of pf_nn_conv_backward_noalloc: one array per
// GPU, loaded with the library's code object and held until the process
// ends. In host code the name `scratch` is only a stand-in for the symbol;
// its device address comes from hipGetSymbolAddress (scratch_on_device).
static __device__ __align__(256) float scratch[pfnn::SCRATCH_FLOATS];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
float* scratch_dev[MAX_DEVICES] = {};   // device address of `scratch`, per device

// The device address of this GPU's scratch array, looked up on the first
// call made with that device current and kept per device.
int scratch_on_device(float** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    snprintf(nn_err, sizeof(nn_err), "pf_nn_conv_backward_noalloc: no usable current device");
    return PF_E_HIP;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      snprintf(nn_err, sizeof(nn_err), "pf_nn_conv_backward_noalloc: hipGetSymbolAddress: %s",
               hipGetErrorString(e));
      return PF_E_HIP;
    }
    scratch_dev[dev] = static_cast<float*>(p);
  }
  *out = scratch_dev[dev];
  return PF_OK;
}

}  // namespace

extern "C" {

int pf_nn_conv_backward_noalloc(const pf_nn_conv* c, const void* x, const void* w, const void* y,
                                const float* dy, float* dx, float* dw, float* dbias, void* stream) {
  if (!valid(c) || !x || !w || !y || !dy)
    return nn_fail("pf_nn_conv_backward_noalloc: bad arguments");
  float* part = nullptr;
  if (dw) {
    const pfnn::Geom g = pfnn::make_geom(*c);
    const int64_t need = (int64_t)pfnn::weight_grad_splits(g) * g.cout * g.kdim;
    if (need > pfnn::SCRATCH_FLOATS) {
      snprintf(nn_err, sizeof(nn_err),
               "pf_nn_conv_backward_noalloc: the weight gradient needs %lld floats of "
               "workspace, more than the %lld-float scratch array",
               (long long)need, (long long)pfnn::SCRATCH_FLOATS);
      return PF_E_LIMIT;
    }
    const int st = scratch_on_device(&part);
    if (st != PF_OK) return st;
  }
  // no BF16 dZ workspace here: the passes round dZ from dY and Y themselves
  // (the same values, so the same bits as pf_nn_conv_backward)
  conv_backward(c, x, w, y, dy, dx, dw, dbias, part, nullptr, static_cast<hipStream_t>(stream));
  return nn_status("pf_nn_conv_backward_noalloc");
}

int pf_nn_mse(int32_t total, const float* pred, const float* target, float* dpred, float* loss,
              void* stream) {
  if (total <= 0 || !pred || !target || !dpred || !loss) return nn_fail("pf_nn_mse: bad arguments");
  const int blocks = (total + 255) / 256 < 256 ? (total + 255) / 256 : 256;
  pfnn::k_mse<<<blocks, 256, 0, static_cast<hipStream_t>(stream)>>>(total, pred, target, dpred,
                                                                       loss);
  return nn_status("pf_nn_mse");
}

int pf_nn_adam(int64_t n, float* param, float* grad, float* m, float* v, void* param_bf16,
               float lr, float beta1, float beta2, float eps, int32_t step, void* stream) {
  if (n <= 0 || !param || !grad || !m || !v || step < 1) return nn_fail("pf_nn_adam: bad arguments");
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  pfnn::k_adam<<<(unsigned)((n + 255) / 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      n, param, grad, m, v, static_cast<uint16_t*>(param_bf16), lr, beta1, beta2, eps, c1, c2);
  return nn_status("pf_nn_adam");
}

int pf_nn_to_bf16(int64_t n, const float* x, void* y, void* stream) {
  if (n <= 0 || !x || !y) return nn_fail("pf_nn_to_bf16: bad arguments");
  pfnn::k_to_bf16<<<(unsigned)((n + 255) / 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      n, x, static_cast<uint16_t*>(y));
  return nn_status("pf_nn_to_bf16");
}

}  // extern "C"
