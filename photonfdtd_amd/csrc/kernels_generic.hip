// Generic Yee + CPML, source, ADE, DFT and flux kernels for gfx950.
//
// These kernels handle every grid shape the scene API allows (1-D, 2-D in any
// plane, 3-D) and are the reference-order ("exact") path: each field update
// performs the operations of the reference NumPy stepper in the same order
// (simulation.py:1252-1309), so a float64 run reproduces the oracle's
// trajectory bit for bit. This file is compiled with -ffp-contract=off.
//
// Thread mapping: blockIdx.y = x index i, blockIdx.x * 256 + tid = flat (j, k)
// index, z fastest, so every wavefront reads 64 consecutive z nodes.
#include "pfhip_internal.h"
#include "yee_common.h"

namespace pfhip {

// PML state of one derivative term; m is the node's compact psi index along
// axis `ax` (negative outside the slab).
template <typename real>
__device__ __forceinline__ void pml_term(Term<real>& t, const StepArgs<real>& a, bool e_side,
                                         int ax, int psi_slot, int i, int j, int k, int m) {
  if (m < 0) {
    t.psi = nullptr;
    return;
  }
  const int node = ax == 0 ? i : (ax == 1 ? j : k);
  const int pc = e_side ? a.pe[ax] : a.ph[ax];
  t.psi = a.psi[psi_slot];
  t.psw = a.psio[psi_slot] ? a.psio[psi_slot] : t.psi;
  t.pid = psi_index(ax, m, i, j, k, a.ny, a.nz, pc);
  t.b = e_side ? a.b_e[ax][node] : a.b_h[ax][node];
  t.c = e_side ? a.c_e[ax][node] : a.c_h[ax][node];
}

// F + K (dA - dB), then F += K psiA', F -= K psiB' on PML nodes, where
// psi' = b psi + c d: the reference's bulk-then-slab order
// (simulation.py:1263-1309).
template <typename real>
__device__ __forceinline__ real apply_terms_v(real f, real K, const Term<real>& A,
                                              const Term<real>& B) {
  real cb;
  if (A.on) {
    cb = A.d;
    if (B.on) cb = cb - B.d;
  } else {
    cb = -B.d;
  }
  f = f + cb * K;
  if (A.on && A.psi) {
    real p = A.b * A.psi[A.pid] + A.c * A.d;
    A.psw[A.pid] = p;
    f = f + K * p;
  }
  if (B.on && B.psi) {
    real p = B.b * B.psi[B.pid] + B.c * B.d;
    B.psw[B.pid] = p;
    f = f - K * p;
  }
  return f;
}

// Every field operand of a node is loaded up front (neighbour indices clamped
// to the node itself where that difference is not taken, so each load is in
// bounds) and the PML maps of the three axes once, so a wavefront keeps all
// of its loads in flight together; the arithmetic is that of the reference
// update, term by term.
template <typename real, bool EXACT>
__global__ __launch_bounds__(256) void k_update_h(StepArgs<real> a) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= a.nyz) return;
  const int i = blockIdx.y;
  int j, k;
  plane_jk(q, a.nyz, a.nz, j, k);
  const int64_t id = (int64_t)i * a.nyz + q;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t sx = a.nyz, sy = nz;
  const real K = -a.ch;
  const real* Ex = a.f[0]; const real* Ey = a.f[1]; const real* Ez = a.f[2];
  const bool xi = nx == 1 || i < nx - 1;
  const bool yj = ny == 1 || j < ny - 1;
  const bool zk = nz == 1 || k < nz - 1;
  const int64_t ixp = nx > 1 && i < nx - 1 ? id + sx : id;
  const int64_t iyp = ny > 1 && j < ny - 1 ? id + sy : id;
  const int64_t izp = nz > 1 && k < nz - 1 ? id + 1 : id;
  const real ex = Ex[id], ey = Ey[id], ez = Ez[id];
  const real ez_y = Ez[iyp], ey_z = Ey[izp], ex_z = Ex[izp];
  const real ez_x = Ez[ixp], ey_x = Ey[ixp], ex_y = Ex[iyp];
  const real hx = a.f[3][id], hy = a.f[4][id], hz = a.f[5][id];
  // out of place: every node is written (unchanged where no term applies)
  const bool oop = a.fo[3] != nullptr;
  real* Hxo = oop ? a.fo[3] : a.f[3];
  real* Hyo = oop ? a.fo[4] : a.f[4];
  real* Hzo = oop ? a.fo[5] : a.f[5];
  real nhx = hx, nhy = hy, nhz = hz;
  bool ux = false, uy = false, uz = false;
  const int mx = nx > 1 ? a.map_h[0][i] : -1;
  const int my = ny > 1 ? a.map_h[1][j] : -1;
  const int mz = nz > 1 ? a.map_h[2][k] : -1;
  Term<real> A, B;
  // Hx -= ch (d_y Ez - d_z Ey)
  if (yj && zk) {
    A.on = ny > 1;
    B.on = nz > 1;
    if (A.on) { A.d = diff<real, EXACT>(ez_y, ez, a.d[1], a.inv_d[1]); pml_term(A, a, false, 1, PSI_HXY, i, j, k, my); }
    if (B.on) { B.d = diff<real, EXACT>(ey_z, ey, a.d[2], a.inv_d[2]); pml_term(B, a, false, 2, PSI_HXZ, i, j, k, mz); }
    if (A.on || B.on) {
      nhx = apply_terms_v(hx, K, A, B);
      ux = true;
    }
  }
  // Hy -= ch (d_z Ex - d_x Ez)
  if (zk && xi) {
    A.on = nz > 1;
    B.on = nx > 1;
    if (A.on) { A.d = diff<real, EXACT>(ex_z, ex, a.d[2], a.inv_d[2]); pml_term(A, a, false, 2, PSI_HYZ, i, j, k, mz); }
    if (B.on) { B.d = diff<real, EXACT>(ez_x, ez, a.d[0], a.inv_d[0]); pml_term(B, a, false, 0, PSI_HYX, i, j, k, mx); }
    if (A.on || B.on) {
      nhy = apply_terms_v(hy, K, A, B);
      uy = true;
    }
  }
  // Hz -= ch (d_x Ey - d_y Ex)
  if (xi && yj) {
    A.on = nx > 1;
    B.on = ny > 1;
    if (A.on) { A.d = diff<real, EXACT>(ey_x, ey, a.d[0], a.inv_d[0]); pml_term(A, a, false, 0, PSI_HZX, i, j, k, mx); }
    if (B.on) { B.d = diff<real, EXACT>(ex_y, ex, a.d[1], a.inv_d[1]); pml_term(B, a, false, 1, PSI_HZY, i, j, k, my); }
    if (A.on || B.on) {
      nhz = apply_terms_v(hz, K, A, B);
      uz = true;
    }
  }
  if (ux || oop) Hxo[id] = nhx;
  if (uy || oop) Hyo[id] = nhy;
  if (uz || oop) Hzo[id] = nhz;
}

template <typename real, bool EXACT>
__global__ __launch_bounds__(256) void k_update_e(StepArgs<real> a) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= a.nyz) return;
  const int i = blockIdx.y;
  int j, k;
  plane_jk(q, a.nyz, a.nz, j, k);
  const int64_t id = (int64_t)i * a.nyz + q;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t sx = a.nyz, sy = nz;
  const bool oop = a.fo[0] != nullptr;
  // H^{n+1/2}: written by the H pass (to fo when out of place)
  const real* Hx = oop ? a.fo[3] : a.f[3];
  const real* Hy = oop ? a.fo[4] : a.f[4];
  const real* Hz = oop ? a.fo[5] : a.f[5];
  real* Exo = oop ? a.fo[0] : a.f[0];
  real* Eyo = oop ? a.fo[1] : a.f[1];
  real* Ezo = oop ? a.fo[2] : a.f[2];
  const bool xi = nx == 1 || i >= 1;
  const bool yj = ny == 1 || j >= 1;
  const bool zk = nz == 1 || k >= 1;
  const int64_t ixm = nx > 1 && i >= 1 ? id - sx : id;
  const int64_t iym = ny > 1 && j >= 1 ? id - sy : id;
  const int64_t izm = nz > 1 && k >= 1 ? id - 1 : id;
  const real hx = Hx[id], hy = Hy[id], hz = Hz[id];
  const real hz_y = Hz[iym], hy_z = Hy[izm], hx_z = Hx[izm];
  const real hz_x = Hz[ixm], hy_x = Hy[ixm], hx_y = Hx[iym];
  const real ex = a.f[0][id], ey = a.f[1][id], ez = a.f[2][id];
  const real K = a.ce[id];
  // subpixel smoothing: each E component has its own coefficient
  const real Ky = a.ce_y ? a.ce_y[id] : K;
  const real Kz = a.ce_z ? a.ce_z[id] : K;
  real nex = ex, ney = ey, nez = ez;
  bool ux = false, uy = false, uz = false;
  const int mx = nx > 1 ? a.map_e[0][i] : -1;
  const int my = ny > 1 ? a.map_e[1][j] : -1;
  const int mz = nz > 1 ? a.map_e[2][k] : -1;
  Term<real> A, B;
  // Ex += ce (d_y Hz - d_z Hy)
  if (yj && zk) {
    A.on = ny > 1;
    B.on = nz > 1;
    if (A.on) { A.d = diff<real, EXACT>(hz, hz_y, a.d[1], a.inv_d[1]); pml_term(A, a, true, 1, PSI_EXY, i, j, k, my); }
    if (B.on) { B.d = diff<real, EXACT>(hy, hy_z, a.d[2], a.inv_d[2]); pml_term(B, a, true, 2, PSI_EXZ, i, j, k, mz); }
    if (A.on || B.on) {
      nex = apply_terms_v(ex, K, A, B);
      ux = true;
    }
  }
  // Ey += ce (d_z Hx - d_x Hz)
  if (zk && xi) {
    A.on = nz > 1;
    B.on = nx > 1;
    if (A.on) { A.d = diff<real, EXACT>(hx, hx_z, a.d[2], a.inv_d[2]); pml_term(A, a, true, 2, PSI_EYZ, i, j, k, mz); }
    if (B.on) { B.d = diff<real, EXACT>(hz, hz_x, a.d[0], a.inv_d[0]); pml_term(B, a, true, 0, PSI_EYX, i, j, k, mx); }
    if (A.on || B.on) {
      ney = apply_terms_v(ey, Ky, A, B);
      uy = true;
    }
  }
  // Ez += ce (d_x Hy - d_y Hx)
  if (xi && yj) {
    A.on = nx > 1;
    B.on = ny > 1;
    if (A.on) { A.d = diff<real, EXACT>(hy, hy_x, a.d[0], a.inv_d[0]); pml_term(A, a, true, 0, PSI_EZX, i, j, k, mx); }
    if (B.on) { B.d = diff<real, EXACT>(hx, hx_y, a.d[1], a.inv_d[1]); pml_term(B, a, true, 1, PSI_EZY, i, j, k, my); }
    if (A.on || B.on) {
      nez = apply_terms_v(ez, Kz, A, B);
      uz = true;
    }
  }
  if (ux || oop) Exo[id] = nex;
  if (uy || oop) Eyo[id] = ney;
  if (uz || oop) Ezo[id] = nez;
}

// Soft sources: one thread per distinct target node; equal targets are
// summed in source order with a rounding to the field dtype after every add,
// which is what `F[i, j, k] += amp * w(t)` does per dipole in the reference
// loop (simulation.py:1366-1390).
template <typename real>
__global__ void k_inject(real* const* __restrict__ fields, const int32_t* __restrict__ g_comp,
                         const int64_t* __restrict__ g_cell, const int32_t* __restrict__ g_begin,
                         const double* __restrict__ amp, const int32_t* __restrict__ wave_row,
                         const double* __restrict__ wave, int64_t n_steps, int64_t step,
                         int32_t n_groups) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_groups) return;
  SourceGroupsDev G;
  G.n_groups = n_groups;
  G.comp = const_cast<int32_t*>(g_comp);
  G.cell = const_cast<int64_t*>(g_cell);
  G.begin = const_cast<int32_t*>(g_begin);
  G.amp = const_cast<double*>(amp);
  G.wave_row = const_cast<int32_t*>(wave_row);
  G.wave = const_cast<double*>(wave);
  inject_group<real>(fields, G, n_steps, step, g);
}

// Batched sources: job blockIdx.y, group blockIdx.x * blockDim.x + lane.
// The step is `step`, or *step_dev when that is set (launches captured in a
// graph read the step from a device counter that k_step_tick advances).
template <typename real>
__global__ void k_src_jobs(const SrcJob<real>* __restrict__ jobs, int64_t step,
                           const int64_t* __restrict__ step_dev) {
  const SrcJob<real>& J = jobs[blockIdx.y];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= J.g.n_groups) return;
  inject_group<real>(J.fields, J.g, J.n_steps, step_dev ? *step_dev : step, g);
}

template <typename real>
void launch_src_jobs(const SrcJob<real>* jobs, int n_jobs, int max_groups, int64_t step, hipStream_t s,
                     const int64_t* step_dev) {
  if (n_jobs <= 0 || max_groups <= 0) return;
  dim3 grid((unsigned)((max_groups + 63) / 64), (unsigned)n_jobs);
  k_src_jobs<real><<<grid, 64, 0, s>>>(jobs, step, step_dev);
}

__global__ void k_step_tick(int64_t* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step = *step + 1;
}
__global__ void k_step_set(int64_t* step, int64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step = v;
}
void launch_step_tick(int64_t* step, hipStream_t s) { k_step_tick<<<1, 64, 0, s>>>(step); }
void launch_step_set(int64_t* step, int64_t v, hipStream_t s) { k_step_set<<<1, 64, 0, s>>>(step, v); }
// Batched DFT capture: job blockIdx.y, element blockIdx.x * blockDim.x +
// lane of [comp][box node], the element order of k_gather_box.
template <typename real>
__global__ void k_dft_jobs(const DftJob<real>* __restrict__ jobs, int64_t step_arg,
                           const int64_t* __restrict__ step_dev) {
  const int64_t step = step_dev ? *step_dev : step_arg;
  const DftJob<real>& J = jobs[blockIdx.y];
  if (step % J.interval != 0) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= J.nbox * J.n_comp) return;
  const int64_t q = step / J.interval;
  const int64_t slot = (J.count0 + (q - J.q0)) % J.K;
  const BoxDev& box = J.box;
  const int c = (int)(t / J.nbox);
  const int64_t r = t - (int64_t)c * J.nbox;
  const int64_t bk = r % box.count[2];
  const int64_t bj = (r / box.count[2]) % box.count[1];
  const int64_t bi = r / (box.count[2] * box.count[1]);
  const int64_t node = (box.start[0] + bi * box.stride[0]) * J.plane +
                       (box.start[1] + bj * box.stride[1]) * J.row +
                       zslot(box.start[2] + bk * box.stride[2], J.zpad);
  J.ring[slot * J.n_comp * J.nbox + t] = J.fields[J.comps[c]][node];
}

template <typename real>
void launch_dft_jobs(const DftJob<real>* jobs, int n_jobs, int64_t max_elems, int64_t step,
                     hipStream_t s, const int64_t* step_dev) {
  if (n_jobs <= 0 || max_elems <= 0) return;
  dim3 grid((unsigned)((max_elems + 255) / 256), (unsigned)n_jobs);
  k_dft_jobs<real><<<grid, 256, 0, s>>>(jobs, step, step_dev);
}
template void launch_dft_jobs<float>(const DftJob<float>*, int, int64_t, int64_t, hipStream_t,
                                     const int64_t*);
template void launch_dft_jobs<double>(const DftJob<double>*, int, int64_t, int64_t, hipStream_t,
                                      const int64_t*);

template void launch_src_jobs<float>(const SrcJob<float>*, int, int, int64_t, hipStream_t,
                                     const int64_t*);
template void launch_src_jobs<double>(const SrcJob<double>*, int, int, int64_t, hipStream_t,
                                      const int64_t*);

// E^n on a dispersive region, captured before the curl update
// (simulation.py:638-643).
template <typename real>
__global__ void k_ade_capture(real* const* __restrict__ fields, const int64_t* __restrict__ cells,
                              int64_t n, real* __restrict__ En) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t c = cells[t];
  En[t] = fields[0][c];
  En[n + t] = fields[1][c];
  En[2 * n + t] = fields[2][c];
}

// Pole recursion and E correction (simulation.py:645-669): per component,
// delta = sum_p (P_p^{n+1} - P_p^n), E -= inv_eps * delta.
template <typename real>
__global__ void k_ade_apply(real* const* __restrict__ fields, const int64_t* __restrict__ cells,
                            int64_t n, const real* __restrict__ En, real* __restrict__ Pcur,
                            real* __restrict__ Pprev, const real* __restrict__ abc, int n_poles,
                            real inv_eps) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || n_poles == 0) return;
  const int64_t cell = cells[t];
  for (int c = 0; c < 3; ++c) {
    const real e = En[c * n + t];
    real delta = 0;
    for (int p = 0; p < n_poles; ++p) {
      const int64_t off = ((int64_t)p * 3 + c) * n + t;
      const real pc = Pcur[off];
      const real pp = Pprev[off];
      const real pn = abc[3 * p] * pc + abc[3 * p + 1] * pp + abc[3 * p + 2] * e;
      const real inc = pn - pc;
      delta = p == 0 ? inc : delta + inc;
      Pprev[off] = pc;
      Pcur[off] = pn;
    }
    real* E = fields[c];
    E[cell] = E[cell] - inv_eps * delta;
  }
}

// Field index of flat box node r (z fastest): the element order of the
// gathered planes.
__device__ inline int64_t box_node(const BoxDev& box, int64_t r, int64_t nbox, int64_t plane,
                                   int row, int zpad) {
  int64_t bi, bj, bk;
  if (nbox <= 0x7FFFFFFF) {
    const uint32_t u = (uint32_t)r, c2 = (uint32_t)box.count[2], c1 = (uint32_t)box.count[1];
    const uint32_t q = u / c2;
    bk = u - q * c2;
    bi = q / c1;
    bj = q - (uint32_t)bi * c1;
  } else {
    const int64_t q = r / box.count[2];
    bk = r - q * box.count[2];
    bi = q / box.count[1];
    bj = q - bi * box.count[1];
  }
  return (box.start[0] + bi * box.stride[0]) * plane + (box.start[1] + bj * box.stride[1]) * row +
         zslot(box.start[2] + bk * box.stride[2], zpad);
}

// Copy a strided box of one or more components into a dense buffer:
// out[(c * cstride) + flat box index] = field_c[node] (cstride = the box's
// node count, or that of a larger box whose x slab this one is).
template <typename real>
__global__ void k_gather_box(real* const* __restrict__ fields, const int32_t* __restrict__ comps,
                             int n_comp, BoxDev box, int64_t plane, int row, int zpad,
                             int64_t cstride, real* __restrict__ out) {
  // blockIdx.y = component (its field pointer is one scalar load per wave);
  // x covers the box's nodes, z fastest
  const int c = blockIdx.y;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nbox = box.count[0] * box.count[1] * box.count[2];
  if (r >= nbox) return;
  const real* F = fields[comps[c]];
  out[(int64_t)c * cstride + r] = F[box_node(box, r, nbox, plane, row, zpad)];
}

// The step's work after the march, in one launch: ADE correction of every
// dispersive medium, the E-side soft sources and the DFT planes' samples.
// The host takes this path only when no source cell is a dispersive cell and
// no sampled E node is one of either, so the three parts touch disjoint E
// nodes and their order does not matter; each part computes exactly what its
// own kernel does (k_ade_jobs, k_inject, k_gather_box). Blocks
// [0, ade_blk[n_ade]) are ADE nodes (job j owns [ade_blk[j], ade_blk[j+1])),
// then src_blk source-group blocks, then the gathers (gather d owns
// [dft_blk[d], dft_blk[d+1]), blocks_c per component).
template <typename real>
__global__ __launch_bounds__(256) void k_post_step(PostArgs<real> P) {
  const int b = (int)blockIdx.x;
  const int tid = (int)threadIdx.x;
  if (b < P.ade_blk[P.n_ade]) {
    int j = 0;
    while (b >= P.ade_blk[j + 1]) ++j;
    const AdeJob<real>& J = P.ade[j];
    const int64_t t = (int64_t)(b - P.ade_blk[j]) * 256 + tid;
    if (t < J.n && J.n_poles > 0)
      ade_pp_node<real>(J.Ein, J.Eout, J.cells, J.n, J.Pcur, J.Pprev, J.abc, J.n_poles, J.inv_eps, t);
    return;
  }
  const int bs = b - P.ade_blk[P.n_ade];
  if (bs < P.src_blk) {
    const int g = bs * 256 + tid;
    if (g < P.src.n_groups) inject_group<real>(P.src_fields, P.src, P.n_steps, P.step, g);
    return;
  }
  const int bd = bs - P.src_blk;
  int d = 0;
  while (d < P.n_dft && bd >= P.dft_blk[d + 1]) ++d;
  if (d >= P.n_dft) return;
  const PostGather<real>& G = P.dft[d];
  const int local = bd - P.dft_blk[d];
  const int c = local / G.blocks_c;
  const int64_t r = (int64_t)(local - c * G.blocks_c) * 256 + tid;
  if (r >= G.nbox) return;
  const real* F = P.dft_fields[G.comps[c]];
  G.out[(int64_t)c * G.nbox + r] = F[box_node(G.box, r, G.nbox, P.plane, P.row, P.zpad)];
}

template <typename real>
void launch_post_step(const PostArgs<real>& P, hipStream_t s) {
  const int blocks = P.ade_blk[P.n_ade] + P.src_blk + P.dft_blk[P.n_dft];
  if (blocks <= 0) return;
  k_post_step<real><<<blocks, 256, 0, s>>>(P);
}

// Fold `count` buffered samples into the complex128 accumulators, one sample
// at a time in step order: acc += (phase * f) * dt, exactly the sequence of
// the reference's per-step `acc += phase * snap * dt` (simulation.py:1133).
// A thread owns one box node of one component for FOLD_FB consecutive
// frequencies (blockIdx.y = component x frequency block, so the phase
// factors are wave-uniform scalar loads), and reads each buffered sample once
// for all of them.
constexpr int FOLD_FB = 8;
constexpr int FOLD_SU = 4;   // samples whose ring loads are in flight together

template <typename real>
__global__ __launch_bounds__(256) void k_dft_fold(const real* __restrict__ ring, int64_t nbox,
                                                  int n_comp, const int32_t* __restrict__ is_h,
                                                  int count, int64_t sample0,
                                                  const double* __restrict__ phase_e,
                                                  const double* __restrict__ phase_h, int n_freq,
                                                  double dt, double* __restrict__ acc) {
  const int nfb = (n_freq + FOLD_FB - 1) / FOLD_FB;
  const int c = (int)blockIdx.y / nfb, f0 = ((int)blockIdx.y % nfb) * FOLD_FB;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nbox) return;
  const int nf = min(FOLD_FB, n_freq - f0);
  const double* ph = is_h[c] ? phase_h : phase_e;
  double* A = acc + 2 * (((int64_t)c * n_freq + f0) * nbox + r);
  double re[FOLD_FB], im[FOLD_FB];
#pragma unroll
  for (int q = 0; q < FOLD_FB; ++q) {
    re[q] = q < nf ? A[2 * q * nbox] : 0.0;
    im[q] = q < nf ? A[2 * q * nbox + 1] : 0.0;
  }
  const real* src = ring + (int64_t)c * nbox + r;
  const int64_t stride = (int64_t)n_comp * nbox;
  // samples in groups of FOLD_SU, their ring loads issued together, then
  // accumulated one sample after another (the per-sample order)
  int s = 0;
  for (; s + FOLD_SU <= count; s += FOLD_SU) {
    double v[FOLD_SU];
#pragma unroll
    for (int u = 0; u < FOLD_SU; ++u) v[u] = (double)src[(int64_t)(s + u) * stride];
#pragma unroll
    for (int u = 0; u < FOLD_SU; ++u) {
      const double* p = ph + 2 * ((sample0 + s + u) * n_freq + f0);
#pragma unroll
      for (int q = 0; q < FOLD_FB; ++q) {
        if (q < nf) {
          re[q] = re[q] + (p[2 * q] * v[u]) * dt;
          im[q] = im[q] + (p[2 * q + 1] * v[u]) * dt;
        }
      }
    }
  }
  for (; s < count; ++s) {
    const double v = (double)src[(int64_t)s * stride];
    const double* p = ph + 2 * ((sample0 + s) * n_freq + f0);
#pragma unroll
    for (int q = 0; q < FOLD_FB; ++q) {
      if (q < nf) {
        re[q] = re[q] + (p[2 * q] * v) * dt;
        im[q] = im[q] + (p[2 * q + 1] * v) * dt;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < FOLD_FB; ++q) {
    if (q < nf) {
      A[2 * q * nbox] = re[q];
      A[2 * q * nbox + 1] = im[q];
    }
  }
}

// Plane Poynting flux, stage 1: per-block partial sums of S . n.
template <typename real>
__global__ __launch_bounds__(256) void k_flux_partial(const real* __restrict__ a1,
                                                      const real* __restrict__ a2,
                                                      const real* __restrict__ b1,
                                                      const real* __restrict__ b2, int axis,
                                                      int64_t index, int nx, int ny, int nz,
                                                      int zrow, int zpad,
                                                      double* __restrict__ partial) {
  __shared__ double red[256];
  const int64_t n1 = axis == 0 ? ny : nx;
  const int64_t n2 = axis == 2 ? ny : nz;
  const int64_t total = n1 * n2;
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int64_t u = t / n2, v = t - (t / n2) * n2;
    int64_t node;
    if (axis == 0) node = (index * ny + u) * zrow + zslot(v, zpad);
    else if (axis == 1) node = (u * ny + index) * zrow + zslot(v, zpad);
    else node = (u * ny + v) * zrow + zslot(index, zpad);
    s += (double)a1[node] * (double)a2[node] - (double)b1[node] * (double)b2[node];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Stage 2: fixed-order sum of the partials, total += (sum * dA) * dt.
__global__ void k_flux_finish(const double* __restrict__ partial, int n_part, double dA, double dt,
                              double* __restrict__ total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int p = 0; p < n_part; ++p) s += partial[p];
  *total = *total + (s * dA) * dt;
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from pfhip.cpp with the handle's stream).
// ---------------------------------------------------------------------------
template <typename real>
void launch_generic_h(const StepArgs<real>& a, bool exact, hipStream_t s) {
  dim3 grid((unsigned)((a.nyz + 255) / 256), (unsigned)a.nx);
  if (exact) k_update_h<real, true><<<grid, 256, 0, s>>>(a);
  else k_update_h<real, false><<<grid, 256, 0, s>>>(a);
}

template <typename real>
void launch_generic_e(const StepArgs<real>& a, bool exact, hipStream_t s) {
  dim3 grid((unsigned)((a.nyz + 255) / 256), (unsigned)a.nx);
  if (exact) k_update_e<real, true><<<grid, 256, 0, s>>>(a);
  else k_update_e<real, false><<<grid, 256, 0, s>>>(a);
}

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename real>
void launch_inject(real* const* fields_dev, const SourceGroupsDev& g, int64_t n_steps, int64_t step,
                   hipStream_t s) {
  if (g.n_groups == 0) return;
  k_inject<real><<<blocks_for(g.n_groups, 64), 64, 0, s>>>(fields_dev, g.comp, g.cell, g.begin,
                                                           g.amp, g.wave_row, g.wave, n_steps, step,
                                                           g.n_groups);
}

template <typename real>
void launch_ade_capture(real* const* fields_dev, const int64_t* cells, int64_t n, real* En,
                        hipStream_t s) {
  k_ade_capture<real><<<blocks_for(n, 256), 256, 0, s>>>(fields_dev, cells, n, En);
}

template <typename real>
void launch_ade_apply(real* const* fields_dev, const int64_t* cells, int64_t n, const real* En,
                      real* Pcur, real* Pprev, const real* abc, int n_poles, real inv_eps,
                      hipStream_t s) {
  k_ade_apply<real><<<blocks_for(n, 256), 256, 0, s>>>(fields_dev, cells, n, En, Pcur, Pprev, abc,
                                                       n_poles, inv_eps);
}

template <typename real>
void launch_gather_box(real* const* fields_dev, const int32_t* comps, int n_comp, const BoxDev& box,
                       int64_t plane, int row, int zpad, real* out, hipStream_t s,
                       int64_t cstride) {
  const int64_t nbox = box.count[0] * box.count[1] * box.count[2];
  if (nbox <= 0 || n_comp <= 0) return;
  const dim3 grid(blocks_for(nbox, 256), (unsigned)n_comp);
  k_gather_box<real><<<grid, 256, 0, s>>>(fields_dev, comps, n_comp, box, plane, row, zpad,
                                          cstride > 0 ? cstride : nbox, out);
}

template <typename real>
void launch_dft_fold(const real* ring, int64_t nbox, int n_comp, const int32_t* is_h, int count,
                     int64_t sample0, const double* phase_e, const double* phase_h, int n_freq,
                     double dt, double* acc, hipStream_t s) {
  if (nbox <= 0 || n_comp <= 0 || n_freq <= 0 || count <= 0) return;
  const dim3 grid(blocks_for(nbox, 256), (unsigned)(n_comp * ((n_freq + FOLD_FB - 1) / FOLD_FB)));
  k_dft_fold<real><<<grid, 256, 0, s>>>(ring, nbox, n_comp, is_h, count, sample0, phase_e,
                                        phase_h, n_freq, dt, acc);
}

template <typename real>
void launch_flux(real* const* fields_host, int axis, int64_t index, int nx, int ny, int nz,
                 int zrow, int zpad, double dA, double dt, double* partial, int n_part,
                 double* total, hipStream_t s) {
  // S . n: x: Ey Hz - Ez Hy; y: Ez Hx - Ex Hz; z: Ex Hy - Ey Hx (simulation.py:1446-1457)
  static const int t1[3][4] = {{1, 5, 2, 4}, {2, 3, 0, 5}, {0, 4, 1, 3}};
  const int* c = t1[axis];
  k_flux_partial<real><<<n_part, 256, 0, s>>>(fields_host[c[0]], fields_host[c[1]],
                                              fields_host[c[2]], fields_host[c[3]], axis, index, nx,
                                              ny, nz, zrow, zpad, partial);
  k_flux_finish<<<1, 64, 0, s>>>(partial, n_part, dA, dt, total);
}

#define PF_INSTANTIATE(real)                                                                       \
  template void launch_generic_h<real>(const StepArgs<real>&, bool, hipStream_t);                  \
  template void launch_generic_e<real>(const StepArgs<real>&, bool, hipStream_t);                  \
  template void launch_inject<real>(real* const*, const SourceGroupsDev&, int64_t, int64_t,        \
                                    hipStream_t);                                                  \
  template void launch_ade_capture<real>(real* const*, const int64_t*, int64_t, real*,             \
                                         hipStream_t);                                             \
  template void launch_ade_apply<real>(real* const*, const int64_t*, int64_t, const real*, real*,  \
                                       real*, const real*, int, real, hipStream_t);                \
  template void launch_gather_box<real>(real* const*, const int32_t*, int, const BoxDev&, int64_t, \
                                        int, int, real*, hipStream_t, int64_t);                    \
  template void launch_post_step<real>(const PostArgs<real>&, hipStream_t);                        \
  template void launch_dft_fold<real>(const real*, int64_t, int, const int32_t*, int, int64_t,     \
                                      const double*, const double*, int, double, double*,          \
                                      hipStream_t);                                                \
  template void launch_flux<real>(real* const*, int, int64_t, int, int, int, int, int, double,     \
                                  double, double*, int, double*, hipStream_t);

PF_INSTANTIATE(float)
PF_INSTANTIATE(double)

}  // namespace pfhip
