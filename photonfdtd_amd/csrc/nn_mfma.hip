// Training and inference kernels of the spectral surrogate on the gfx950
// matrix cores.
//
// Every convolution pass is an implicit GEMM on v_mfma_f32_16x16x16_bf16
// (BF16 operands, FP32 accumulation): a workgroup of four wavefronts owns a
// 64x64 tile of the output matrix, stages 64x16 slices of both operands in
// LDS per K step (gathered straight from the NHWC activations, so no im2col
// buffer exists), and each wavefront issues 2x2 MFMAs on its 32x32 quarter.
//
//   forward      Y[m = pixel][n = cout]   = sum_k X~[m][k = tap, cin] W[n][k]
//   data grad    dX[m = pixel][n = cin]   = sum_k dZ~[m][k = tap, cout] W[cout][tap][n]
//   weight grad  dW[m = cout][n = k]      = sum_r dZ[r][m] X~[r][n]   (split over r)
//
// X~ / dZ~ are the gathered (im2col / transposed-im2col) operands;
// dZ = dY * relu'(Y) is formed during the gather. Activations are stored as
// BF16 (raw 16-bit, round to nearest even), gradients as FP32 and rounded to
// BF16 when they enter an MFMA. Weights are kept in FP32 by the optimizer with
// a BF16 copy for the MFMAs.
//
// Operand lane layout of v_mfma_f32_16x16x16_bf16 (checked on the device):
// lane l holds A[l % 16][4 (l / 16) + j] and B[4 (l / 16) + j][l % 16],
// j = 0..3, and receives D[4 (l / 16) + j][l % 16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <cstdint>
#include <cstdio>

#include "../../include/pfhip.h"

namespace pfnn {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;   // NaN
  u += 0x7fffu + ((u >> 16) & 1u);                        // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

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

// im2col element: output pixel r of the convolution, k = (tap, cin).
__device__ __forceinline__ uint16_t im2col(const Geom& g, const uint16_t* x, int r, int kk) {
  const int ci = kk % g.cin, tap = kk / g.cin;
  const int ky = tap / g.k, kx = tap - ky * g.k;
  const int ox = r % g.wo, t = r / g.wo;
  const int oy = t % g.ho, b = t / g.ho;
  const int iy = oy * g.s + ky - g.pad, ix = ox * g.s + kx - g.pad;
  if (iy < 0 || iy >= g.h || ix < 0 || ix >= g.w) return 0;
  return x[((size_t)(b * g.h + iy) * g.w + ix) * g.cin + ci];
}

// dZ = dY * relu'(Y) at output pixel r, channel co (Y is BF16 or FP32).
__device__ __forceinline__ float dz(const Geom& g, const float* dy, const void* y, int relu,
                                    int out_f32, int r, int co) {
  const size_t i = (size_t)r * g.cout + co;
  float v = dy[i];
  if (relu) {
    const float yv = out_f32 ? static_cast<const float*>(y)[i]
                             : bf2f(static_cast<const uint16_t*>(y)[i]);
    if (!(yv > 0.f)) v = 0.f;
  }
  return v;
}

struct FwdOp {
  Geom g;
  const uint16_t* x;
  const uint16_t* w;
  const float* bias;
  void* y;
  int relu, out_f32;
  __device__ uint16_t a(int m, int kk) const { return im2col(g, x, m, kk); }
  __device__ uint16_t b(int kk, int nn) const { return w[(size_t)nn * g.kdim + kk]; }
  __device__ void store(int m, int nn, float acc) const {
    float v = acc + (bias ? bias[nn] : 0.f);
    if (relu && !(v > 0.f)) v = 0.f;
    const size_t i = (size_t)m * g.cout + nn;
    if (out_f32) static_cast<float*>(y)[i] = v;
    else static_cast<uint16_t*>(y)[i] = f2bf(v);
  }
};

struct DataGradOp {   // M = input pixels, N = cin, K = (tap, cout)
  Geom g;
  const uint16_t* w;
  const float* dy;
  const void* y;
  int relu, out_f32;
  float* dx;
  __device__ uint16_t a(int m, int kk) const {
    const int co = kk % g.cout, tap = kk / g.cout;
    const int ky = tap / g.k, kx = tap - ky * g.k;
    const int ix = m % g.w, t = m / g.w;
    const int iy = t % g.h, b = t / g.h;
    const int ny = iy + g.pad - ky, nx = ix + g.pad - kx;
    if (ny < 0 || nx < 0 || ny % g.s || nx % g.s) return 0;
    const int oy = ny / g.s, ox = nx / g.s;
    if (oy >= g.ho || ox >= g.wo) return 0;
    return f2bf(dz(g, dy, y, relu, out_f32, (b * g.ho + oy) * g.wo + ox, co));
  }
  __device__ uint16_t b(int kk, int nn) const {
    const int co = kk % g.cout, tap = kk / g.cout;
    return w[((size_t)co * g.k * g.k + tap) * g.cin + nn];
  }
  __device__ void store(int m, int nn, float acc) const { dx[(size_t)m * g.cin + nn] = acc; }
};

struct WeightGradOp {   // M = cout, N = kdim, K = output pixels (split)
  Geom g;
  const uint16_t* x;
  const float* dy;
  const void* y;
  int relu, out_f32;
  float* part;          // [split][cout][kdim]: the partial sum of each pixel split
  __device__ uint16_t a(int m, int r) const { return f2bf(dz(g, dy, y, relu, out_f32, r, m)); }
  __device__ uint16_t b(int r, int nn) const { return im2col(g, x, r, nn); }
  __device__ void store(int m, int nn, float acc) const {
    part[((size_t)blockIdx.z * g.cout + m) * g.kdim + nn] = acc;
  }
};

// dw[i] += sum over the splits of part[z][i], z = 0, 1, ... in order: the
// weight gradient is the same bits on every run (no atomics).
__global__ void k_reduce_splits(const float* part, int splits, int64_t count, float* dw) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = part[i];
  for (int z = 1; z < splits; ++z) s += part[(size_t)z * count + i];
  dw[i] += s;
}

constexpr int TM = 64, TN = 64, TK = 16, LDS_ROW = TK + 4;

// C[M x N] (+)= A[M x K] B[K x N] over K range [z * kper, (z + 1) * kper).
template <class Op>
__global__ __launch_bounds__(256) void k_gemm(Op op, int M, int N, int K, int kper) {
  __shared__ uint16_t sA[TM][LDS_ROW];
  __shared__ uint16_t sB[TN][LDS_ROW];   // stored as [n][k]
  const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
  const int kb = blockIdx.z * kper, ke = min(K, kb + kper);
  const int t = threadIdx.x, l = t & 63, wv = t >> 6;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  const int row = t >> 2, kq = (t & 3) * 4;
  f4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += TK) {
    for (int j = 0; j < 4; ++j) {
      const int kk = k0 + kq + j;
      sA[row][kq + j] = (m0 + row < M && kk < ke) ? op.a(m0 + row, kk) : (uint16_t)0;
      sB[row][kq + j] = (n0 + row < N && kk < ke) ? op.b(kk, n0 + row) : (uint16_t)0;
    }
    __syncthreads();
    s4 fa[2], fb[2];
    for (int i = 0; i < 2; ++i) {
      fa[i] = *reinterpret_cast<const s4*>(&sA[wm + 16 * i + (l & 15)][4 * (l >> 4)]);
      fb[i] = *reinterpret_cast<const s4*>(&sB[wn + 16 * i + (l & 15)][4 * (l >> 4)]);
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[i], fb[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (l >> 4) + r;
        const int nn = n0 + wn + 16 * j + (l & 15);
        if (m < M && nn < N) op.store(m, nn, acc[i][j][r]);
      }
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
// kper = ceil(kk / splits) rounded up to TK.
inline int split_count(int kk, int splits) {
  const int kper = ((kk + splits - 1) / splits + TK - 1) / TK * TK;
  return (kk + kper - 1) / kper;
}

template <class Op>
void launch_gemm(const Op& op, int M, int N, int K, int splits, hipStream_t s) {
  const int kper = ((K + splits - 1) / splits + TK - 1) / TK * TK;
  const int z = (K + kper - 1) / kper;
  dim3 grid((M + TM - 1) / TM, (N + TN - 1) / TN, z);
  k_gemm<Op><<<grid, 256, 0, s>>>(op, M, N, K, kper);
}

// Largest weight-gradient workspace a conv backward may need from the
// library's static scratch array, in floats (16 MiB); the pixel splits of a
// layer are capped so that splits * cout * kdim stays within it.
constexpr int64_t SCRATCH_FLOATS = int64_t(4) << 20;

// Pixel splits of the weight-gradient GEMM of a layer: about 4096 output
// pixels each, at most SCRATCH_FLOATS / (cout kdim) (at least 1). A function
// of the shape only, so both backward entry points sum the same partials in
// the same order.
inline int weight_grad_splits(const Geom& g) {
  const int64_t R = (int64_t)g.n * g.ho * g.wo;
  const int64_t per = (int64_t)g.cout * g.kdim;
  int64_t s = R > 4096 ? (R + 4095) / 4096 : 1;
  s = std::min<int64_t>(s, std::max<int64_t>(1, SCRATCH_FLOATS / per));
  return split_count((int)R, (int)s);
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
  pfnn::FwdOp op{g, static_cast<const uint16_t*>(x), static_cast<const uint16_t*>(w), bias, y,
                 c->relu, c->out_f32};
  pfnn::launch_gemm(op, g.n * g.ho * g.wo, g.cout, g.kdim, 1, static_cast<hipStream_t>(stream));
  return nn_status("pf_nn_conv_forward");
}

}  // extern "C"

namespace {

// The backward kernels of one layer; `part` is the weight-gradient workspace
// of weight_grad_splits(g) * cout * kdim floats (unused when dw is NULL).
void conv_backward(const pf_nn_conv* c, const void* x, const void* w, const void* y,
                   const float* dy, float* dx, float* dw, float* dbias, float* part,
                   hipStream_t s) {
  const pfnn::Geom g = pfnn::make_geom(*c);
  if (dx) {
    pfnn::DataGradOp op{g, static_cast<const uint16_t*>(w), dy, y, c->relu, c->out_f32, dx};
    pfnn::launch_gemm(op, g.n * g.h * g.w, g.cin, g.k * g.k * g.cout, 1, s);
  }
  if (dw) {
    const int splits = pfnn::weight_grad_splits(g);
    pfnn::WeightGradOp op{g, static_cast<const uint16_t*>(x), dy, y, c->relu, c->out_f32, part};
    pfnn::launch_gemm(op, g.cout, g.kdim, g.n * g.ho * g.wo, splits, s);
    const int64_t count = (int64_t)g.cout * g.kdim;
    pfnn::k_reduce_splits<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(part, splits, count, dw);
  }
  if (dbias) pfnn::k_bias_grad<<<g.cout, 256, 0, s>>>(g, dy, y, c->relu, c->out_f32, dbias);
}

}  // namespace

extern "C" {

int pf_nn_conv_backward(const pf_nn_conv* c, const void* x, const void* w, const void* y,
                        const float* dy, float* dx, float* dw, float* dbias, void* stream) {
  if (!valid(c) || !x || !w || !y || !dy) return nn_fail("pf_nn_conv_backward: bad arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  float* part = nullptr;
  if (dw) {
    // stream-ordered workspace: allocated, used and released in order on s
    const pfnn::Geom g = pfnn::make_geom(*c);
    const size_t bytes = (size_t)pfnn::weight_grad_splits(g) * g.cout * g.kdim * sizeof(float);
    if (hipMallocAsync(reinterpret_cast<void**>(&part), bytes, s) != hipSuccess)
      return nn_status("pf_nn_conv_backward: workspace allocation");
  }
  conv_backward(c, x, w, y, dy, dx, dw, dbias, part, s);
  const int st = nn_status("pf_nn_conv_backward");
  if (part) (void)hipFreeAsync(part, s);
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
  conv_backward(c, x, w, y, dy, dx, dw, dbias, part, static_cast<hipStream_t>(stream));
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
