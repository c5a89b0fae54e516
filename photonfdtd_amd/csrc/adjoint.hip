// Discrete adjoint of the Yee + CPML step (the transpose of the generic
// in-place update) for gfx950.
//
// Forward step (per node, simulation.py:1252-1309): for a component F with
// terms A (+) and B (-) on difference operators D_A, D_B,
//   dA = D_A(src_A), psiA' = bA psiA + cA dA   (CPML nodes only)
//   F' = F + K (dA - dB + psiA' - psiB'),  K = ce (E) or -dt/mu0 (H).
// Its transpose, given lam = dJ/dF' and the running adjoints of psi:
//   g = K lam
//   LA = lam_psiA + g,  lam_psiA <- bA LA,  lam_dA = g + cA LA   (else lam_dA = g)
//   LB = lam_psiB - g,  lam_psiB <- bB LB,  lam_dB = -g + cB LB  (else lam_dB = -g)
//   lam_src += D^T lam_d ;  dJ/dce += lam (dA - dB + psiA' - psiB')   (E only)
// The per-term lam_d are written to scratch arrays by one kernel and gathered
// onto the source components by the next (D^T is a two-point stencil): the
// E update's terms at the start of the H transpose of the same reverse step,
// the H update's terms at the start of the E transpose of the next one.
#include "pfhip_internal.h"

namespace pfhip {

namespace {

__device__ __forceinline__ int64_t adj_psi_index(int ax, int m, int i, int j, int k, int ny,
                                                 int nz, int pc) {
  if (ax == 0) return ((int64_t)m * ny + j) * nz + k;
  if (ax == 1) return ((int64_t)i * pc + m) * nz + k;
  return ((int64_t)i * ny + j) * pc + m;
}

// One term's transpose. Returns lam_d and adds the forward term (d + psi')
// times `sign` to `curl`.
// m is the node's compact psi index along axis ax (negative outside the slab).
template <typename real>
__device__ __forceinline__ real term_t(const AdjArgs<real>& a, bool e_side, int ax, int slot, int i,
                                       int j, int k, int m, real g, real sign, real d, real& curl) {
  if (m < 0) {
    curl += sign * d;
    return sign * g;
  }
  const int node = ax == 0 ? i : (ax == 1 ? j : k);
  const int pc = e_side ? a.pe[ax] : a.ph[ax];
  const int64_t pid = adj_psi_index(ax, m, i, j, k, a.ny, a.nz, pc);
  const real b = (e_side ? a.b_e[ax] : a.b_h[ax])[node];
  const real c = (e_side ? a.c_e[ax] : a.c_h[ax])[node];
  const real psi_new = b * a.psi_fwd[slot][pid] + c * d;
  curl += sign * (d + psi_new);
  const real L = a.lam_psi[slot][pid] + sign * g;
  a.lam_psi[slot][pid] = b * L;
  return sign * g + c * L;
}

// lam_H(id) += D^T lam_d for the backward differences of the E update:
// (D^T v)[q] = (v[q] - v[q + s]) / d, v[q + s] = 0 past the last node.
template <typename real>
__device__ __forceinline__ void gather_e_terms(const AdjArgs<real>& a, int64_t id, int i, int j,
                                               int k, real& hx, real& hy, real& hz) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t sx = a.nyz, sy = nz;
  const bool nxt_x = i + 1 < nx, nxt_y = j + 1 < ny, nxt_z = k + 1 < nz;
  const int64_t ixp = nxt_x ? id + sx : id, iyp = nxt_y ? id + sy : id, izp = nxt_z ? id + 1 : id;
  auto dt_ = [&](real v, real vn, bool has_next, real inv) {
    return (v - (has_next ? vn : real(0))) * inv;
  };
  const real *D0 = a.lam_d[0], *D1 = a.lam_d[1], *D2 = a.lam_d[2];
  const real *D3 = a.lam_d[3], *D4 = a.lam_d[4], *D5 = a.lam_d[5];
  const real d2 = D2[id], d2n = D2[izp], d5 = D5[id], d5n = D5[iyp];
  const real d1 = D1[id], d1n = D1[izp], d4 = D4[id], d4n = D4[ixp];
  const real d0 = D0[id], d0n = D0[iyp], d3 = D3[id], d3n = D3[ixp];
  if (nz > 1) hx += dt_(d2, d2n, nxt_z, a.inv_d[2]);      // Ey: +d_z Hx
  if (ny > 1) hx += dt_(d5, d5n, nxt_y, a.inv_d[1]);      // Ez: -d_y Hx
  if (nz > 1) hy += dt_(d1, d1n, nxt_z, a.inv_d[2]);      // Ex: -d_z Hy
  if (nx > 1) hy += dt_(d4, d4n, nxt_x, a.inv_d[0]);      // Ez: +d_x Hy
  if (ny > 1) hz += dt_(d0, d0n, nxt_y, a.inv_d[1]);      // Ex: +d_y Hz
  if (nx > 1) hz += dt_(d3, d3n, nxt_x, a.inv_d[0]);      // Ey: -d_x Hz
}

// lam_E(id) += D^T lam_dh for the forward differences of the H update:
// (D^T v)[q] = (v[q - s] - v[q]) / d, v[q - s] = 0 before the first node.
template <typename real>
__device__ __forceinline__ void gather_h_terms(const AdjArgs<real>& a, int64_t id, int i, int j,
                                               int k, real& ex, real& ey, real& ez) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t sx = a.nyz, sy = nz;
  const bool prv_x = i >= 1, prv_y = j >= 1, prv_z = k >= 1;
  const int64_t ixm = prv_x ? id - sx : id, iym = prv_y ? id - sy : id, izm = prv_z ? id - 1 : id;
  auto dt_ = [&](real v, real vp, bool has_prev, real inv) {
    return ((has_prev ? vp : real(0)) - v) * inv;
  };
  const real *D0 = a.lam_dh[0], *D1 = a.lam_dh[1], *D2 = a.lam_dh[2];
  const real *D3 = a.lam_dh[3], *D4 = a.lam_dh[4], *D5 = a.lam_dh[5];
  const real d2 = D2[id], d2p = D2[izm], d5 = D5[id], d5p = D5[iym];
  const real d1 = D1[id], d1p = D1[izm], d4 = D4[id], d4p = D4[ixm];
  const real d0 = D0[id], d0p = D0[iym], d3 = D3[id], d3p = D3[ixm];
  if (nz > 1) ex += dt_(d2, d2p, prv_z, a.inv_d[2]);      // Hy: +d_z Ex
  if (ny > 1) ex += dt_(d5, d5p, prv_y, a.inv_d[1]);      // Hz: -d_y Ex
  if (nz > 1) ey += dt_(d1, d1p, prv_z, a.inv_d[2]);      // Hx: -d_z Ey
  if (nx > 1) ey += dt_(d4, d4p, prv_x, a.inv_d[0]);      // Hz: +d_x Ey
  if (ny > 1) ez += dt_(d0, d0p, prv_y, a.inv_d[1]);      // Hx: +d_y Ez
  if (nx > 1) ez += dt_(d3, d3p, prv_x, a.inv_d[0]);      // Hy: -d_x Ez
}

}  // namespace

// The two per-node kernels below load every operand of a node up front:
// neighbour indices are clamped to the node itself where a difference is not
// taken (so each load is in bounds) and the PML maps of the three axes are
// read once, so a wavefront keeps all of its loads in flight together.

// Transpose of the E update. Forward inputs: H^{n+1/2} (a.f[3..5]) and the
// E-side psi before the update (a.psi_fwd[0..5]).
template <typename real>
__global__ __launch_bounds__(256) void k_adj_e(AdjArgs<real> a, int gather) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= a.nyz) return;
  const int i = blockIdx.y;
  int j, k;
  plane_jk(q, a.nyz, a.nz, j, k);
  const int64_t id = (int64_t)i * a.nyz + q;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t sx = a.nyz, sy = nz;
  const real* Hx = a.f[3]; const real* Hy = a.f[4]; const real* Hz = a.f[5];
  const bool xi = nx == 1 || i >= 1, yj = ny == 1 || j >= 1, zk = nz == 1 || k >= 1;
  const int64_t ixm = nx > 1 && i >= 1 ? id - sx : id;
  const int64_t iym = ny > 1 && j >= 1 ? id - sy : id;
  const int64_t izm = nz > 1 && k >= 1 ? id - 1 : id;
  const real hx = Hx[id], hy = Hy[id], hz = Hz[id];
  const real hz_y = Hz[iym], hy_z = Hy[izm], hx_z = Hx[izm];
  const real hz_x = Hz[ixm], hy_x = Hy[ixm], hx_y = Hx[iym];
  real l0 = a.lam_f[0][id], l1 = a.lam_f[1][id], l2 = a.lam_f[2][id];
  if (gather) {
    // the H-update terms of the reverse step before (step n + 1)
    gather_h_terms(a, id, i, j, k, l0, l1, l2);
    a.lam_f[0][id] = l0;
    a.lam_f[1][id] = l1;
    a.lam_f[2][id] = l2;
  }
  const real K = a.ce[id];
  const double g0 = a.grad[id];
  const int mx = nx > 1 ? a.map_e[0][i] : -1;
  const int my = ny > 1 ? a.map_e[1][j] : -1;
  const int mz = nz > 1 ? a.map_e[2][k] : -1;
  double gsum = 0.0;
  // slots of lam_d: 0 ExY 1 ExZ 2 EyZ 3 EyX 4 EzX 5 EzY
  {  // Ex: +d_y Hz, -d_z Hy
    real dA_l = 0, dB_l = 0, curl = 0;
    if (yj && zk) {
      const real lam = l0, g = K * lam;
      if (ny > 1) dA_l = term_t(a, true, 1, PSI_EXY, i, j, k, my, g, real(1), (hz - hz_y) * a.inv_d[1], curl);
      if (nz > 1) dB_l = term_t(a, true, 2, PSI_EXZ, i, j, k, mz, g, real(-1), (hy - hy_z) * a.inv_d[2], curl);
      gsum += (double)lam * (double)curl;
    }
    a.lam_d[0][id] = dA_l;
    a.lam_d[1][id] = dB_l;
  }
  {  // Ey: +d_z Hx, -d_x Hz
    real dA_l = 0, dB_l = 0, curl = 0;
    if (zk && xi) {
      const real lam = l1, g = K * lam;
      if (nz > 1) dA_l = term_t(a, true, 2, PSI_EYZ, i, j, k, mz, g, real(1), (hx - hx_z) * a.inv_d[2], curl);
      if (nx > 1) dB_l = term_t(a, true, 0, PSI_EYX, i, j, k, mx, g, real(-1), (hz - hz_x) * a.inv_d[0], curl);
      gsum += (double)lam * (double)curl;
    }
    a.lam_d[2][id] = dA_l;
    a.lam_d[3][id] = dB_l;
  }
  {  // Ez: +d_x Hy, -d_y Hx
    real dA_l = 0, dB_l = 0, curl = 0;
    if (xi && yj) {
      const real lam = l2, g = K * lam;
      if (nx > 1) dA_l = term_t(a, true, 0, PSI_EZX, i, j, k, mx, g, real(1), (hy - hy_x) * a.inv_d[0], curl);
      if (ny > 1) dB_l = term_t(a, true, 1, PSI_EZY, i, j, k, my, g, real(-1), (hx - hx_y) * a.inv_d[1], curl);
      gsum += (double)lam * (double)curl;
    }
    a.lam_d[4][id] = dA_l;
    a.lam_d[5][id] = dB_l;
  }
  a.grad[id] = g0 + gsum;
}

// Transpose of the H update. Forward inputs: E^n (a.f[0..2]) and the H-side
// psi before the update (a.psi_fwd[6..11]).
template <typename real>
__global__ __launch_bounds__(256) void k_adj_h(AdjArgs<real> a) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= a.nyz) return;
  const int i = blockIdx.y;
  int j, k;
  plane_jk(q, a.nyz, a.nz, j, k);
  const int64_t id = (int64_t)i * a.nyz + q;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t sx = a.nyz, sy = nz;
  const real* Ex = a.f[0]; const real* Ey = a.f[1]; const real* Ez = a.f[2];
  const real K = -a.ch;
  const bool xi = nx == 1 || i < nx - 1, yj = ny == 1 || j < ny - 1, zk = nz == 1 || k < nz - 1;
  const int64_t ixp = nx > 1 && i < nx - 1 ? id + sx : id;
  const int64_t iyp = ny > 1 && j < ny - 1 ? id + sy : id;
  const int64_t izp = nz > 1 && k < nz - 1 ? id + 1 : id;
  const real ex = Ex[id], ey = Ey[id], ez = Ez[id];
  const real ez_y = Ez[iyp], ey_z = Ey[izp], ex_z = Ex[izp];
  const real ez_x = Ez[ixp], ey_x = Ey[ixp], ex_y = Ex[iyp];
  // the E-update terms of this reverse step first
  real l3 = a.lam_f[3][id], l4 = a.lam_f[4][id], l5 = a.lam_f[5][id];
  gather_e_terms(a, id, i, j, k, l3, l4, l5);
  a.lam_f[3][id] = l3;
  a.lam_f[4][id] = l4;
  a.lam_f[5][id] = l5;
  const int mx = nx > 1 ? a.map_h[0][i] : -1;
  const int my = ny > 1 ? a.map_h[1][j] : -1;
  const int mz = nz > 1 ? a.map_h[2][k] : -1;
  real unused = 0;
  // slots of lam_d: 0 HxY 1 HxZ 2 HyZ 3 HyX 4 HzX 5 HzY
  {  // Hx: +d_y Ez, -d_z Ey
    real dA_l = 0, dB_l = 0;
    if (yj && zk) {
      const real g = K * l3;
      if (ny > 1) dA_l = term_t(a, false, 1, PSI_HXY, i, j, k, my, g, real(1), (ez_y - ez) * a.inv_d[1], unused);
      if (nz > 1) dB_l = term_t(a, false, 2, PSI_HXZ, i, j, k, mz, g, real(-1), (ey_z - ey) * a.inv_d[2], unused);
    }
    a.lam_dh[0][id] = dA_l;
    a.lam_dh[1][id] = dB_l;
  }
  {  // Hy: +d_z Ex, -d_x Ez
    real dA_l = 0, dB_l = 0;
    if (zk && xi) {
      const real g = K * l4;
      if (nz > 1) dA_l = term_t(a, false, 2, PSI_HYZ, i, j, k, mz, g, real(1), (ex_z - ex) * a.inv_d[2], unused);
      if (nx > 1) dB_l = term_t(a, false, 0, PSI_HYX, i, j, k, mx, g, real(-1), (ez_x - ez) * a.inv_d[0], unused);
    }
    a.lam_dh[2][id] = dA_l;
    a.lam_dh[3][id] = dB_l;
  }
  {  // Hz: +d_x Ey, -d_y Ex
    real dA_l = 0, dB_l = 0;
    if (xi && yj) {
      const real g = K * l5;
      if (nx > 1) dA_l = term_t(a, false, 0, PSI_HZX, i, j, k, mx, g, real(1), (ey_x - ey) * a.inv_d[0], unused);
      if (ny > 1) dB_l = term_t(a, false, 1, PSI_HZY, i, j, k, my, g, real(-1), (ex_y - ex) * a.inv_d[1], unused);
    }
    a.lam_dh[4][id] = dA_l;
    a.lam_dh[5][id] = dB_l;
  }
}

// Adjoint source of a DFT monitor at sample q: for each captured node,
// lam += dt * sum_f Re(conj(G_f) phase_{q,f}), where G = dJ/d(Re F) + i dJ/d(Im F).
template <typename real>
__global__ void k_adj_dft_source(real* const* __restrict__ lam, const int32_t* __restrict__ comps,
                                 const int32_t* __restrict__ is_h, int n_comp, BoxDev box,
                                 int64_t nyz, int nz, const double* __restrict__ cot, int n_freq,
                                 const double* __restrict__ phase_e,
                                 const double* __restrict__ phase_h, int64_t sample, double dt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nbox = box.count[0] * box.count[1] * box.count[2];
  if (t >= nbox * n_comp) return;
  const int c = (int)(t / nbox);
  const int64_t r = t - (int64_t)c * nbox;
  const double* ph = (is_h[c] ? phase_h : phase_e) + 2 * sample * n_freq;
  const double* G = cot + (int64_t)c * n_freq * nbox * 2;
  double s = 0.0;
  for (int f = 0; f < n_freq; ++f) {
    const double* g = G + 2 * ((int64_t)f * nbox + r);
    s += g[0] * ph[2 * f] + g[1] * ph[2 * f + 1];
  }
  const int64_t bk = r % box.count[2];
  const int64_t bj = (r / box.count[2]) % box.count[1];
  const int64_t bi = r / (box.count[2] * box.count[1]);
  const int64_t node = (box.start[0] + bi * box.stride[0]) * nyz +
                       (box.start[1] + bj * box.stride[1]) * nz + (box.start[2] + bk * box.stride[2]);
  real* L = lam[comps[c]];
  L[node] = (real)((double)L[node] + s * dt);
}

template <typename real>
void launch_adjoint_step(const AdjArgs<real>& a, hipStream_t s, int phase, bool gather) {
  dim3 grid((unsigned)((a.nyz + 255) / 256), (unsigned)a.nx);
  if (phase == 0)
    k_adj_e<real><<<grid, 256, 0, s>>>(a, gather ? 1 : 0);
  else
    k_adj_h<real><<<grid, 256, 0, s>>>(a);
}

template <typename real>
void launch_adj_dft_source(real* const* lam_dev, const int32_t* comps, const int32_t* is_h,
                           int n_comp, const BoxDev& box, int64_t nyz, int nz, const double* cot,
                           int n_freq, const double* phase_e, const double* phase_h,
                           int64_t sample, double dt, hipStream_t s) {
  const int64_t nbox = box.count[0] * box.count[1] * box.count[2];
  const int64_t n = nbox * n_comp;
  k_adj_dft_source<real><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(
      lam_dev, comps, is_h, n_comp, box, nyz, nz, cot, n_freq, phase_e, phase_h, sample, dt);
}

template void launch_adjoint_step<float>(const AdjArgs<float>&, hipStream_t, int, bool);
template void launch_adjoint_step<double>(const AdjArgs<double>&, hipStream_t, int, bool);
template void launch_adj_dft_source<float>(float* const*, const int32_t*, const int32_t*, int,
                                           const BoxDev&, int64_t, int, const double*, int,
                                           const double*, const double*, int64_t, double,
                                           hipStream_t);
template void launch_adj_dft_source<double>(double* const*, const int32_t*, const int32_t*, int,
                                            const BoxDev&, int64_t, int, const double*, int,
                                            const double*, const double*, int64_t, double,
                                            hipStream_t);

}  // namespace pfhip
