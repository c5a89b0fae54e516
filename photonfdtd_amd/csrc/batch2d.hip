// Batched 2-D stepping: many independent xy-plane simulations of different
// sizes advanced together by one kernel per phase, or by one persistent
// cooperative kernel that loops over all time steps with a grid barrier
// between the H and E phases.
//
// Only the active polarisation group is stepped (TE: Ex, Ey, Hz; TM: Ez, Hx,
// Hy): in 2-D the two groups decouple, so with sources in one group the other
// stays exactly zero, as in the reference's six-component update. The owning
// thread of a node also applies the node's soft sources, its ADE pole update
// (E phase) and its DFT accumulation, in the reference's order (curl update,
// ADE, sources, then the DFT sample; simulation.py:1319-1406), with the same
// per-node arithmetic as the generic kernels.
#include <hip/hip_cooperative_groups.h>

#include "pfhip_internal.h"
#include "yee_common.h"

namespace pfhip {

namespace {

template <typename real>
__device__ __forceinline__ void psi_term(Term<real>& t, const SimDev<real>& s, bool e_side, int ax,
                                         int slot, int i, int j) {
  const int node = ax == 0 ? i : j;
  const int m = (e_side ? s.map_e[ax] : s.map_h[ax])[node];
  if (m < 0) {
    t.psi = nullptr;
    return;
  }
  const int pc = e_side ? s.pe[ax] : s.ph[ax];
  t.psi = s.psi[slot];
  t.pid = ax == 0 ? (int64_t)m * s.ny + j : (int64_t)i * pc + m;
  t.b = (e_side ? s.b_e[ax] : s.b_h[ax])[node];
  t.c = (e_side ? s.c_e[ax] : s.c_h[ax])[node];
}

// Soft sources of one side landing on (comp, cell), in source order.
template <typename real>
__device__ __forceinline__ real add_sources(const SimDev<real>& s, int side, int comp, int64_t cell,
                                            int64_t step, real v) {
  if (cell < s.src_lo[side] || cell > s.src_hi[side]) return v;
  const SourceGroupsDev& g = side ? s.src_h : s.src_e;
  const int ng = g.n_groups;
  for (int q = 0; q < ng; ++q) {
    if (g.comp[q] != comp || g.cell[q] != cell) continue;
    for (int t = g.begin[q]; t < g.begin[q + 1]; ++t)
      v = (real)((double)v + g.amp[t] * g.wave[(int64_t)g.wave_row[t] * s.n_steps + step]);
  }
  return v;
}

// DFT sample of one component value at a node, if the node is in a monitor box.
template <typename real>
__device__ __forceinline__ void dft_sample(const SimDev<real>& s, int comp, int i, int j,
                                           int64_t step, real v) {
  for (int m = 0; m < s.n_dft; ++m) {
    const DftDev& d = s.dft[m];
    const int di = i - d.lo[0], dj = j - d.lo[1];
    if (di < 0 || dj < 0 || i > d.hi[0] || j > d.hi[1]) continue;
    if (d.interval > 1 && step % d.interval != 0) continue;
    int slot = -1;
    for (int c = 0; c < d.n_comp; ++c)
      if (d.comps[c] == comp) slot = c;
    if (slot < 0) continue;
    const int sx = (int)d.box.stride[0], sy = (int)d.box.stride[1];
    if ((sx > 1 && di % sx) || (sy > 1 && dj % sy)) continue;
    const int64_t bi = sx > 1 ? di / sx : di, bj = sy > 1 ? dj / sy : dj;
    const int64_t r = bi * d.box.count[1] + bj;
    const int64_t nbox = d.box.count[0] * d.box.count[1];
    const int64_t q = step / d.interval;
    const double* ph = (comp >= 3 ? d.phase_h : d.phase_e) + 2 * q * d.n_freq;
    double* A = d.acc + 2 * ((int64_t)slot * d.n_freq * nbox + r);
    const double x = (double)v;
    for (int f = 0; f < d.n_freq; ++f) {
      double* a = A + 2 * (int64_t)f * nbox;
      a[0] = a[0] + (ph[2 * f] * x) * d.dt;
      a[1] = a[1] + (ph[2 * f + 1] * x) * d.dt;
    }
  }
}

// ADE for an E component at a node (simulation.py:645-669); en = E^n.
template <typename real>
__device__ __forceinline__ real ade_update(const SimDev<real>& s, int code, int comp, real en,
                                           real e) {
  const int med = code >> 26, t = code & ((1 << 26) - 1);
  const AdeDev<real>& M = s.ade[med];
  if (M.n_poles == 0) return e;
  real delta = 0;
  for (int p = 0; p < M.n_poles; ++p) {
    const int64_t off = ((int64_t)p * 3 + comp) * M.n + t;
    const real pc = M.Pcur[off], pp = M.Pprev[off];
    const real pn = M.abc[3 * p] * pc + M.abc[3 * p + 1] * pp + M.abc[3 * p + 2] * en;
    const real inc = pn - pc;
    delta = p == 0 ? inc : delta + inc;
    M.Pprev[off] = pc;
    M.Pcur[off] = pn;
  }
  return e - M.inv_eps * delta;
}

template <typename real, bool EXACT>
__device__ __forceinline__ void node_h(const SimDev<real>& s, int id, int64_t step) {
  const int nx = s.nx, ny = s.ny;
  const int i = id / ny, j = id - i * ny;
  const real K = -s.ch;
  Term<real> A, B;
  if (s.pol == 0) {   // TE: Hz -= ch (d_x Ey - d_y Ex)
    if (i < nx - 1 && j < ny - 1) {
      A.on = true;
      B.on = true;
      A.d = diff<real, EXACT>(s.f[1][id + ny], s.f[1][id], s.d[0], s.inv_d[0]);
      psi_term(A, s, false, 0, PSI_HZX, i, j);
      B.d = diff<real, EXACT>(s.f[0][id + 1], s.f[0][id], s.d[1], s.inv_d[1]);
      psi_term(B, s, false, 1, PSI_HZY, i, j);
      apply_terms(s.f[5], id, K, A, B);
    }
    real v = add_sources(s, 1, 5, id, step, s.f[5][id]);
    s.f[5][id] = v;
    dft_sample(s, 5, i, j, step, v);
  } else {            // TM: Hx -= ch d_y Ez ; Hy -= ch (- d_x Ez)
    B.on = false;
    if (j < ny - 1) {
      A.on = true;
      A.d = diff<real, EXACT>(s.f[2][id + 1], s.f[2][id], s.d[1], s.inv_d[1]);
      psi_term(A, s, false, 1, PSI_HXY, i, j);
      apply_terms(s.f[3], id, K, A, B);
    }
    A.on = false;
    if (i < nx - 1) {
      B.on = true;
      B.d = diff<real, EXACT>(s.f[2][id + ny], s.f[2][id], s.d[0], s.inv_d[0]);
      psi_term(B, s, false, 0, PSI_HYX, i, j);
      apply_terms(s.f[4], id, K, A, B);
    }
    for (int c = 3; c <= 4; ++c) {
      real v = add_sources(s, 1, c, id, step, s.f[c][id]);
      s.f[c][id] = v;
      dft_sample(s, c, i, j, step, v);
    }
  }
}

template <typename real, bool EXACT>
__device__ __forceinline__ void node_e(const SimDev<real>& s, int id, int64_t step) {
  const int ny = s.ny;
  const int i = id / ny, j = id - i * ny;
  const real K = s.ce[id];
  const int code = s.ade_code ? s.ade_code[id] : -1;
  Term<real> A, B;
  if (s.pol == 0) {   // TE: Ex += ce d_y Hz ; Ey += ce (- d_x Hz)
    const real ex0 = s.f[0][id], ey0 = s.f[1][id];
    B.on = false;
    if (j >= 1) {
      A.on = true;
      A.d = diff<real, EXACT>(s.f[5][id], s.f[5][id - 1], s.d[1], s.inv_d[1]);
      psi_term(A, s, true, 1, PSI_EXY, i, j);
      apply_terms(s.f[0], id, K, A, B);
    }
    A.on = false;
    if (i >= 1) {
      B.on = true;
      B.d = diff<real, EXACT>(s.f[5][id], s.f[5][id - ny], s.d[0], s.inv_d[0]);
      psi_term(B, s, true, 0, PSI_EYX, i, j);
      apply_terms(s.f[1], id, K, A, B);
    }
    real ex = s.f[0][id], ey = s.f[1][id];
    if (code >= 0) {
      ex = ade_update(s, code, 0, ex0, ex);
      ey = ade_update(s, code, 1, ey0, ey);   // the idle Ez poles stay zero
    }
    ex = add_sources(s, 0, 0, id, step, ex);
    ey = add_sources(s, 0, 1, id, step, ey);
    s.f[0][id] = ex;
    s.f[1][id] = ey;
    dft_sample(s, 0, i, j, step, ex);
    dft_sample(s, 1, i, j, step, ey);
  } else {            // TM: Ez += ce (d_x Hy - d_y Hx)
    const real ez0 = s.f[2][id];
    if (i >= 1 && j >= 1) {
      A.on = true;
      B.on = true;
      A.d = diff<real, EXACT>(s.f[4][id], s.f[4][id - ny], s.d[0], s.inv_d[0]);
      psi_term(A, s, true, 0, PSI_EZX, i, j);
      B.d = diff<real, EXACT>(s.f[3][id], s.f[3][id - 1], s.d[1], s.inv_d[1]);
      psi_term(B, s, true, 1, PSI_EZY, i, j);
      apply_terms(s.f[2], id, K, A, B);
    }
    real ez = s.f[2][id];
    if (code >= 0) ez = ade_update(s, code, 2, ez0, ez);   // idle Ex/Ey poles stay zero
    ez = add_sources(s, 0, 2, id, step, ez);
    s.f[2][id] = ez;
    dft_sample(s, 2, i, j, step, ez);
  }
}

// One phase (0 = H, 1 = E) over the work chunks owned by this block.
template <typename real, bool EXACT>
__device__ __forceinline__ void run_phase(const BatchArgs<real>& a, int phase, int64_t step) {
  for (int ch = blockIdx.x; ch < a.n_chunks; ch += gridDim.x) {
    const int sim = a.chunk_sim[ch];
    const SimDev<real>& s = a.sims[sim];
    if (step >= s.n_steps) continue;
    const int c0 = a.chunk_first[ch];
    const int c1 = min(c0 + a.chunk_cells, s.nx * s.ny);
    for (int id = c0 + (int)threadIdx.x; id < c1; id += (int)blockDim.x) {
      if (phase == 0) node_h<real, EXACT>(s, id, step);
      else node_e<real, EXACT>(s, id, step);
    }
  }
}

}  // namespace

// Per-phase kernel; the step is read from a device counter so a captured
// graph of several steps can be replayed (k_batch_tick advances it).
template <typename real, bool EXACT>
__global__ void k_batch_phase(BatchArgs<real> a, int phase) {
  run_phase<real, EXACT>(a, phase, *a.step);
}

__global__ void k_batch_tick(int64_t* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step = *step + 1;
}

// Persistent kernel: all steps [step0, step1) in one cooperative launch.
template <typename real, bool EXACT>
__global__ void k_batch_persistent(BatchArgs<real> a, int64_t step0, int64_t step1) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  for (int64_t st = step0; st < step1; ++st) {
    run_phase<real, EXACT>(a, 0, st);
    grid.sync();
    run_phase<real, EXACT>(a, 1, st);
    grid.sync();
  }
}

template <typename real>
void launch_batch_phase(const BatchArgs<real>& a, int phase, bool exact, int grid, int block,
                        hipStream_t s) {
  if (exact) k_batch_phase<real, true><<<grid, block, 0, s>>>(a, phase);
  else k_batch_phase<real, false><<<grid, block, 0, s>>>(a, phase);
}

void launch_batch_tick(int64_t* step, hipStream_t s) { k_batch_tick<<<1, 64, 0, s>>>(step); }

// Workgroups of the persistent kernel: every workgroup must be resident at
// once for the grid barrier, so the grid is capped at PERSISTENT_MAX_PER_CU
// workgroups per CU (below what the occupancy calculator allows) and at the
// number of work chunks.
constexpr int PERSISTENT_MAX_PER_CU = 4;

template <typename real>
hipError_t batch_persistent_grid(bool exact, int block, int n_chunks, int* per_cu, int* n_cu,
                                 int* grid) {
  const void* fn = exact ? reinterpret_cast<const void*>(&k_batch_persistent<real, true>)
                         : reinterpret_cast<const void*>(&k_batch_persistent<real, false>);
  int dev = 0;
  *per_cu = 0;
  *n_cu = 0;
  *grid = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fn, block, 0);
  if (e != hipSuccess) return e;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  if (*per_cu < 1 || *n_cu < 1) return hipErrorInvalidConfiguration;
  *grid = std::max(1, std::min(std::min(*per_cu, PERSISTENT_MAX_PER_CU) * *n_cu, n_chunks));
  return hipSuccess;
}

template <typename real>
hipError_t launch_batch_persistent(BatchArgs<real>& a, bool exact, int block, int64_t step0,
                                   int64_t step1, hipStream_t s) {
  int per_cu = 0, n_cu = 0, grid = 0;
  hipError_t e = batch_persistent_grid<real>(exact, block, a.n_chunks, &per_cu, &n_cu, &grid);
  if (e != hipSuccess) return e;
  const void* fn = exact ? reinterpret_cast<const void*>(&k_batch_persistent<real, true>)
                         : reinterpret_cast<const void*>(&k_batch_persistent<real, false>);
  void* args[] = {&a, &step0, &step1};
  return hipLaunchCooperativeKernel(fn, dim3(grid), dim3(block), args, 0, s);
}

template void launch_batch_phase<float>(const BatchArgs<float>&, int, bool, int, int, hipStream_t);
template void launch_batch_phase<double>(const BatchArgs<double>&, int, bool, int, int, hipStream_t);
template hipError_t launch_batch_persistent<float>(BatchArgs<float>&, bool, int, int64_t, int64_t,
                                                   hipStream_t);
template hipError_t launch_batch_persistent<double>(BatchArgs<double>&, bool, int, int64_t, int64_t,
                                                    hipStream_t);
template hipError_t batch_persistent_grid<float>(bool, int, int, int*, int*, int*);
template hipError_t batch_persistent_grid<double>(bool, int, int, int*, int*, int*);

}  // namespace pfhip
