// Batched 2-D stepping: many independent xy-plane simulations of different
// sizes advanced by one march kernel per time step, or by one persistent
// kernel that loops over the steps with a grid barrier between them.
//
// Each wavefront owns a strip of 62 columns (lanes 1..62 of a 64-lane
// window) over a run of rows and marches down the rows. At row i it computes
// H^{n+1/2}(i) from E^n(i) (registers) and E^n(i+1) (loaded one row ahead),
// then E^{n+1}(i) from H^{n+1/2}(i) and H^{n+1/2}(i-1) (registers). The
// neighbour columns come from the adjacent lanes (lane 63 supplies E(j+1),
// lane 0 recomputes H(j-1)); the first row recomputes H(i0-1). Fields and
// H-side psi are double-buffered (read level p = step & 1, write p ^ 1), so
// the recomputed halo reads only old values; every node is written by one
// lane. Per node and step this reads E (2 components), H (1) and ce once and
// writes E and H once: 28 bytes in float32, the Yee minimum.
//
// Only the polarisation group the sources excite is stepped (TE: Ex, Ey, Hz;
// TM: Ez, Hx, Hy): in 2-D the two groups decouple and the other one stays
// exactly zero, as in the reference's six-component update. The march does
// the curl updates with the CPML slab terms and the H-side soft sources; a
// post pass over the few nodes that need more then applies, per node and in
// the reference's order (simulation.py:1319-1406), the ADE correction, the
// E-side soft sources and the DFT sample. Every node does the arithmetic of
// the generic kernels, so a batched simulation reproduces its single run bit
// for bit.
#include "pfhip_internal.h"
#include "yee_common.h"

namespace pfhip {

// Field rows are stored with the streaming hint: a row is read again only in
// the next step, and the lines it would push out of the caches hold the
// apron columns that the neighbouring strips read.
template <typename real>
__device__ __forceinline__ void st_row(real* p, real v) {
  __builtin_nontemporal_store(v, p);
}

namespace {

// CPML membership of an axis: node n is in the slab set iff n < lo or
// n >= hi; m(n) is its compact psi index.
struct Slab {
  int lo, hi;
  __device__ __forceinline__ bool in(int n) const { return n < lo || n >= hi; }
  __device__ __forceinline__ int m(int n) const { return n < lo ? n : lo + (n - hi); }
};

// Soft sources of one group (one component and cell), in source order
// (simulation.py:1366-1390): one rounding to the field type per add.
template <typename real>
__device__ __forceinline__ real add_group(const SourceGroupsDev& g, int q, int64_t n_steps,
                                          int64_t step, real v) {
  for (int t = g.begin[q]; t < g.begin[q + 1]; ++t)
    v = (real)((double)v + g.amp[t] * g.wave[(int64_t)g.wave_row[t] * n_steps + step]);
  return v;
}

template <typename real>
__device__ __forceinline__ real add_h_sources(const SimDev<real>& s, int comp, int64_t cell,
                                              int64_t step, real v) {
  if (cell < s.src_lo || cell > s.src_hi) return v;
  const SourceGroupsDev& g = s.src_h;
  for (int q = 0; q < g.n_groups; ++q)
    if (g.comp[q] == comp && g.cell[q] == cell) v = add_group(g, q, s.n_steps, step, v);
  return v;
}

// ADE of one E component at a node (simulation.py:645-669): en = E^n,
// e = curl-updated E^{n+1}; P^{n+1} = a P^n + b P^{n-1} + c E^n and
// E -= inv_eps * sum_p (P^{n+1} - P^n).
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

// ---- TE: Hz -= ch (d_x Ey - d_y Ex); Ex += ce d_y Hz; Ey += ce (-d_x Hz) ----
template <typename real>
struct RowTE {
  real ex, ey;       // E^n(i + 1)
  real hz, ce;       // H^{n-1/2}(i), ce(i)
  real phx, phy;     // Hz psi: x term (HZX), y term (HZY), level p
  real pey, pex;     // Ex psi y term (EXY), Ey psi x term (EYX)
};

// Field rows of a batched simulation are padded: column j sits in slot
// zslot(j, MARCH_OWNED), so each 62-column strip owns an aligned 64-slot
// block. A lane loads its column's slot; it stores into the strip's block
// (lanes 1..62 their columns, lanes 0 and 63 the block's two pad slots), so
// a row store covers whole lines.
__device__ __forceinline__ int strip_load_slot(int jc) { return (int)zslot(jc, MARCH_OWNED); }
__device__ __forceinline__ int strip_store_slot(const MarchChunk& c, int lane) {
  return (c.j0 + 1) / MARCH_OWNED * ZPAD_BLOCK +
         (lane == 0 ? MARCH_OWNED : lane == 63 ? MARCH_OWNED + 1 : lane - 1);
}

template <typename real, bool EXACT, bool HSRC>
__device__ __forceinline__ void march_te(const SimDev<real>& s, const MarchChunk& c, int p,
                                         int64_t step) {
  const int lane = threadIdx.x & 63;
  const int nx = s.nx, ny = s.ny;
  const int j = c.j0 + lane;
  const bool col = j >= 0 && j < ny;
  const bool own = col && lane >= 1 && lane <= MARCH_OWNED;
  const int jc = min(max(j, 0), ny - 1);
  const real* Ex = s.f[p][0];
  const real* Ey = s.f[p][1];
  const real* Hz = s.f[p][5];
  real* Exo = s.f[p ^ 1][0];
  real* Eyo = s.f[p ^ 1][1];
  real* Hzo = s.f[p ^ 1][5];
  const real* hin_x = s.psiH[p][PSI_HZX - 6];
  const real* hin_y = s.psiH[p][PSI_HZY - 6];
  real* hout_x = s.psiH[p ^ 1][PSI_HZX - 6];
  real* hout_y = s.psiH[p ^ 1][PSI_HZY - 6];
  real* pe_y = s.psiE[PSI_EXY];
  real* pe_x = s.psiE[PSI_EYX];
  const Slab xe{s.slab_e[0][0], s.slab_e[0][1]}, xh{s.slab_h[0][0], s.slab_h[0][1]};
  const Slab ye{s.slab_e[1][0], s.slab_e[1][1]}, yh{s.slab_h[1][0], s.slab_h[1][1]};
  const bool pyh = col && yh.in(jc), pye = own && ye.in(jc);
  const int myh = yh.m(jc), mye = ye.m(jc);
  const real byh = pyh ? s.b_h[1][jc] : real(0), cyh = pyh ? s.c_h[1][jc] : real(0);
  const real bye = pye ? s.b_e[1][jc] : real(0), cye = pye ? s.c_e[1][jc] : real(0);
  const int ph1 = s.ph[1], pe1 = s.pe[1];
  const real d0 = s.d[0], d1 = s.d[1], r0 = s.inv_d[0], r1 = s.inv_d[1];
  const real Kh = -s.ch;
  const int i0 = c.i0, i1 = c.i1, ib = i0 > 0 ? i0 - 1 : 0;
  const int nyr = s.nyr, sj = strip_load_slot(jc), ss = strip_store_slot(c, lane);

  auto fetch = [&](RowTE<real>& R, int i) {
    const int o = i * nyr + sj;
    R.ex = R.ey = real(0);
    if (col && i + 1 < nx) {
      R.ex = Ex[o + nyr];
      R.ey = Ey[o + nyr];
    }
    R.hz = col ? Hz[o] : real(0);
    R.ce = own ? s.ce[o] : real(0);
    R.phx = (col && xh.in(i)) ? hin_x[xh.m(i) * ny + jc] : real(0);
    R.phy = pyh ? hin_y[i * ph1 + myh] : real(0);
    R.pey = pye ? pe_y[i * pe1 + mye] : real(0);
    R.pex = (own && xe.in(i)) ? pe_x[xe.m(i) * ny + jc] : real(0);
  };

  real ex_c = real(0), ey_c = real(0);   // E^n(i)
  if (col) {
    ex_c = Ex[ib * nyr + sj];
    ey_c = Ey[ib * nyr + sj];
  }
  real hz_prev = real(0);                // H^{n+1/2}(i - 1)
  auto row = [&](int i, const RowTE<real>& R, RowTE<real>& N) {
    if (i + 1 < i1) fetch(N, i + 1);
    const bool live = i >= i0;
    const real ex_r = __shfl_down(ex_c, 1);
    // H^{n+1/2}(i)
    const bool pxh = xh.in(i);
    const bool hupd = col && i < nx - 1 && j < ny - 1;
    real hz = R.hz, nhx = real(0), nhy = real(0);
    if (hupd) {
      const real bxh = pxh ? s.b_h[0][i] : real(0), cxh = pxh ? s.c_h[0][i] : real(0);
      const real dA = diff<real, EXACT>(R.ey, ey_c, d0, r0);
      const real dB = diff<real, EXACT>(ex_r, ex_c, d1, r1);
      hz = curl_update(hz, Kh, true, dA, pxh, R.phx, bxh, cxh, &nhx, true, dB, pyh, R.phy, byh,
                       cyh, &nhy);
    }
    if (HSRC && col) hz = add_h_sources(s, 5, (int64_t)i * ny + j, step, hz);
    if (live) st_row(&Hzo[i * nyr + ss], hz);   // every lane: the row store covers whole lines
    if (own && live) {
      if (hupd && pxh) hout_x[xh.m(i) * ny + j] = nhx;
      if (hupd && pyh) hout_y[i * ph1 + myh] = nhy;
    }
    const real hz_l = __shfl_up(hz, 1);
    // E^{n+1}(i)
    real ex = ex_c, ey = ey_c;
    if (own && live) {
      const real K = R.ce;
      const bool pxe = xe.in(i);
      real nex = real(0), ney = real(0), unused = real(0);
      if (j >= 1) {
        const real dA = diff<real, EXACT>(hz, hz_l, d1, r1);
        ex = curl_update(ex, K, true, dA, pye, R.pey, bye, cye, &nex, false, real(0), false,
                         real(0), real(0), real(0), &unused);
      }
      if (i >= 1) {
        const real bxe = pxe ? s.b_e[0][i] : real(0), cxe = pxe ? s.c_e[0][i] : real(0);
        const real dB = diff<real, EXACT>(hz, hz_prev, d0, r0);
        ey = curl_update(ey, K, false, real(0), false, real(0), real(0), real(0), &unused, true,
                         dB, pxe, R.pex, bxe, cxe, &ney);
      }
      if (j >= 1 && pye) pe_y[i * pe1 + mye] = nex;
      if (i >= 1 && pxe) pe_x[xe.m(i) * ny + j] = ney;
    }
    if (live) {
      st_row(&Exo[i * nyr + ss], ex);
      st_row(&Eyo[i * nyr + ss], ey);
    }
    hz_prev = hz;
  };
  // Two row buffers alternate so no register still being loaded is copied.
  RowTE<real> A, B;
  fetch(A, ib);
  for (int i = ib; i < i1; i += 2) {
    if (i > ib) {
      ex_c = B.ex;
      ey_c = B.ey;
    }
    row(i, A, B);
    if (i + 1 >= i1) break;
    ex_c = A.ex;
    ey_c = A.ey;
    row(i + 1, B, A);
  }
}

// ---- TM: Hx -= ch d_y Ez; Hy -= ch (-d_x Ez); Ez += ce (d_x Hy - d_y Hx) ----
template <typename real>
struct RowTM {
  real ez;           // E^n(i + 1)
  real hx, hy, ce;   // H^{n-1/2}(i), ce(i)
  real phy, phx;     // Hx psi y term (HXY), Hy psi x term (HYX), level p
  real pex, pey;     // Ez psi x term (EZX), y term (EZY)
};

template <typename real, bool EXACT, bool HSRC>
__device__ __forceinline__ void march_tm(const SimDev<real>& s, const MarchChunk& c, int p,
                                         int64_t step) {
  const int lane = threadIdx.x & 63;
  const int nx = s.nx, ny = s.ny;
  const int j = c.j0 + lane;
  const bool col = j >= 0 && j < ny;
  const bool own = col && lane >= 1 && lane <= MARCH_OWNED;
  const int jc = min(max(j, 0), ny - 1);
  const real* Ez = s.f[p][2];
  const real* Hx = s.f[p][3];
  const real* Hy = s.f[p][4];
  real* Ezo = s.f[p ^ 1][2];
  real* Hxo = s.f[p ^ 1][3];
  real* Hyo = s.f[p ^ 1][4];
  const real* hin_y = s.psiH[p][PSI_HXY - 6];
  const real* hin_x = s.psiH[p][PSI_HYX - 6];
  real* hout_y = s.psiH[p ^ 1][PSI_HXY - 6];
  real* hout_x = s.psiH[p ^ 1][PSI_HYX - 6];
  real* pe_x = s.psiE[PSI_EZX];
  real* pe_y = s.psiE[PSI_EZY];
  const Slab xe{s.slab_e[0][0], s.slab_e[0][1]}, xh{s.slab_h[0][0], s.slab_h[0][1]};
  const Slab ye{s.slab_e[1][0], s.slab_e[1][1]}, yh{s.slab_h[1][0], s.slab_h[1][1]};
  const bool pyh = col && yh.in(jc), pye = own && ye.in(jc);
  const int myh = yh.m(jc), mye = ye.m(jc);
  const real byh = pyh ? s.b_h[1][jc] : real(0), cyh = pyh ? s.c_h[1][jc] : real(0);
  const real bye = pye ? s.b_e[1][jc] : real(0), cye = pye ? s.c_e[1][jc] : real(0);
  const int ph1 = s.ph[1], pe1 = s.pe[1];
  const real d0 = s.d[0], d1 = s.d[1], r0 = s.inv_d[0], r1 = s.inv_d[1];
  const real Kh = -s.ch;
  const int i0 = c.i0, i1 = c.i1, ib = i0 > 0 ? i0 - 1 : 0;
  const int nyr = s.nyr, sj = strip_load_slot(jc), ss = strip_store_slot(c, lane);

  auto fetch = [&](RowTM<real>& R, int i) {
    const int o = i * nyr + sj;
    R.ez = (col && i + 1 < nx) ? Ez[o + nyr] : real(0);
    R.hx = col ? Hx[o] : real(0);
    R.hy = col ? Hy[o] : real(0);
    R.ce = own ? s.ce[o] : real(0);
    R.phy = pyh ? hin_y[i * ph1 + myh] : real(0);
    R.phx = (col && xh.in(i)) ? hin_x[xh.m(i) * ny + jc] : real(0);
    R.pex = (own && xe.in(i)) ? pe_x[xe.m(i) * ny + jc] : real(0);
    R.pey = pye ? pe_y[i * pe1 + mye] : real(0);
  };

  real ez_c = col ? Ez[ib * nyr + sj] : real(0);   // E^n(i)
  real hy_prev = real(0);                          // Hy^{n+1/2}(i - 1)
  auto row = [&](int i, const RowTM<real>& R, RowTM<real>& N) {
    if (i + 1 < i1) fetch(N, i + 1);
    const bool live = i >= i0;
    const real ez_r = __shfl_down(ez_c, 1);
    const bool pxh = xh.in(i);
    real hx = R.hx, hy = R.hy, nhy = real(0), nhx = real(0), unused = real(0);
    const bool xupd = col && j < ny - 1, yupd = col && i < nx - 1;
    if (xupd) {
      const real dA = diff<real, EXACT>(ez_r, ez_c, d1, r1);
      hx = curl_update(hx, Kh, true, dA, pyh, R.phy, byh, cyh, &nhy, false, real(0), false,
                       real(0), real(0), real(0), &unused);
    }
    if (yupd) {
      const real bxh = pxh ? s.b_h[0][i] : real(0), cxh = pxh ? s.c_h[0][i] : real(0);
      const real dB = diff<real, EXACT>(R.ez, ez_c, d0, r0);
      hy = curl_update(hy, Kh, false, real(0), false, real(0), real(0), real(0), &unused, true, dB,
                       pxh, R.phx, bxh, cxh, &nhx);
    }
    if (HSRC && col) {
      const int64_t cell = (int64_t)i * ny + j;
      hx = add_h_sources(s, 3, cell, step, hx);
      hy = add_h_sources(s, 4, cell, step, hy);
    }
    if (live) {   // every lane: the row store covers whole lines
      st_row(&Hxo[i * nyr + ss], hx);
      st_row(&Hyo[i * nyr + ss], hy);
    }
    if (own && live) {
      if (xupd && pyh) hout_y[i * ph1 + myh] = nhy;
      if (yupd && pxh) hout_x[xh.m(i) * ny + j] = nhx;
    }
    const real hx_l = __shfl_up(hx, 1);
    real ez = ez_c;
    if (own && live) {
      const real K = R.ce;
      const bool pxe = xe.in(i);
      real nzx = real(0), nzy = real(0);
      const bool eupd = i >= 1 && j >= 1;
      if (eupd) {
        const real bxe = pxe ? s.b_e[0][i] : real(0), cxe = pxe ? s.c_e[0][i] : real(0);
        const real dA = diff<real, EXACT>(hy, hy_prev, d0, r0);
        const real dB = diff<real, EXACT>(hx, hx_l, d1, r1);
        ez = curl_update(ez, K, true, dA, pxe, R.pex, bxe, cxe, &nzx, true, dB, pye, R.pey, bye,
                         cye, &nzy);
      }
      if (eupd && pxe) pe_x[xe.m(i) * ny + j] = nzx;
      if (eupd && pye) pe_y[i * pe1 + mye] = nzy;
    }
    if (live) st_row(&Ezo[i * nyr + ss], ez);
    hy_prev = hy;
  };
  RowTM<real> A, B;
  fetch(A, ib);
  for (int i = ib; i < i1; i += 2) {
    if (i > ib) ez_c = B.ez;
    row(i, A, B);
    if (i + 1 >= i1) break;
    ez_c = A.ez;
    row(i + 1, B, A);
  }
}

template <typename real, bool EXACT, int POL, bool HSRC>
__device__ __forceinline__ void march_chunk(const BatchArgs<real>& a, const MarchChunk& c,
                                            int64_t step) {
  const SimDev<real>& s = a.sims[c.sim];
  if (step >= s.n_steps) return;
  const int p = (int)(step & 1);
  if (POL == 0) march_te<real, EXACT, HSRC>(s, c, p, step);
  else march_tm<real, EXACT, HSRC>(s, c, p, step);
}

// Fold the DFT samples of steps [s_lo, s_hi) (clamped to the simulation's
// own step count) of entry e at frequency f from the ring into its
// accumulator: acc += (phase * v) * dt per sample in step order
// (simulation.py:1116-1133).
template <typename real>
__device__ void dft_fold(const BatchArgs<real>& a, int e, int f, int64_t s_lo, int64_t s_hi) {
  const SimDev<real>& s = a.sims[a.dft_sim[e]];
  const DftEntry D = a.dft[e];
  const DftDev& d = s.dft[D.mon];
  if (f >= d.n_freq) return;
  const int64_t hi = min(s_hi, s.n_steps);
  if (hi <= s_lo) return;
  const int64_t iv = d.interval;
  const int64_t q0 = (s_lo + iv - 1) / iv, q1 = (hi - 1) / iv;
  if (q1 < q0) return;
  const double* ph = (D.comp >= 3 ? d.phase_h : d.phase_e) + 2 * f;
  double* acc = d.acc + 2 * (((int64_t)D.slot * d.n_freq + f) * d.nbox + D.r);
  double re = acc[0], im = acc[1];
  for (int64_t q = q0; q <= q1; ++q) {
    const double v = (double)a.ring[(q % BATCH_DFT_RING) * a.n_dft + e];
    const double* pq = ph + 2 * q * d.n_freq;
    re = re + (pq[0] * v) * d.dt;
    im = im + (pq[1] * v) * d.dt;
  }
  acc[0] = re;
  acc[1] = im;
}

// Post pass of one node: ADE correction of the curl-updated E (E^n from
// level p), E-side soft sources, then the DFT samples of the node.
template <typename real>
__device__ void post_node(const BatchArgs<real>& a, int idx, int64_t step) {
  const PostEntry P = a.post[idx];
  const SimDev<real>& s = a.sims[P.sim];
  if (step >= s.n_steps) return;
  const int p = (int)(step & 1), po = p ^ 1;
  const int cell = P.cell;
  if (P.ade_code >= 0) {
    const int c0 = s.pol == 0 ? 0 : 2, c1 = s.pol == 0 ? 1 : 2;   // idle poles stay zero
    for (int c = c0; c <= c1; ++c) {
      real* E = s.f[po][c];
      E[cell] = ade_update(s, P.ade_code, c, s.f[p][c][cell], E[cell]);
    }
  }
  for (int k = P.src0; k < P.src1; ++k) {
    const int q = a.src_list[k];
    real* F = s.f[po][s.src_e.comp[q]];
    F[cell] = add_group(s.src_e, q, s.n_steps, step, F[cell]);
  }
  for (int e = P.dft0; e < P.dft1; ++e) {
    const DftEntry D = a.dft[e];
    const int interval = s.dft[D.mon].interval;
    if (interval > 1 && step % interval != 0) continue;
    const int64_t q = step / interval;
    a.ring[(q % BATCH_DFT_RING) * a.n_dft + e] = s.f[po][D.comp][cell];
  }
}

__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
}

// Grid barrier with a bounded wait. The last workgroup to arrive resets the
// arrival count and advances the generation; the others wait for the
// generation to change. A wait that outlasts wait_ticks of the wall clock
// (workgroups that are not all resident at once) sets *fail, and every
// workgroup then leaves its step loop, so the kernel always finishes.
__device__ bool grid_barrier(unsigned* bar, int* fail, uint64_t wait_ticks) {
  __shared__ int ok_shared;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned n = gridDim.x;
    const unsigned gen = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned arrived =
        __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1;
    if (arrived == n) {
      __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const uint64_t t0 = wall_clock64();
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        if (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          ok = 0;
          break;
        }
        if (wall_clock64() - t0 > wait_ticks) {
          __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    ok_shared = ok;
  }
  __syncthreads();
  return ok_shared != 0;
}

}  // namespace

// One time step of the chunks [first, first + count) (one polarisation and
// H-source variant); the step is read from a device counter so a captured
// graph of several steps can be replayed (the post pass advances it).
template <typename real, bool EXACT, int POL, bool HSRC>
__global__ __launch_bounds__(512) void k_batch_march(BatchArgs<real> a, int first, int count) {
  const int w = wave_id();
  if (w >= count) return;
  march_chunk<real, EXACT, POL, HSRC>(a, a.chunks[first + w], *a.step);
}

// Post pass of one step. The last workgroup to finish advances the device
// step counter (every workgroup has read it by then), so no separate launch
// is needed for it.
template <typename real>
__global__ __launch_bounds__(256) void k_batch_post(BatchArgs<real> a) {
  __shared__ int last;
  const int64_t step = *a.step;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < a.n_post) post_node(a, idx, step);
  // Every thread of this workgroup has read the counter before the barrier,
  // and every workgroup counts itself only after that, so the last one may
  // advance it. The next launch sees the new value at the kernel boundary,
  // so relaxed atomics suffice (no cache write-back per workgroup).
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned done =
        __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = done == gridDim.x - 1;
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __hip_atomic_store(a.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a.step, step + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Fold the ring samples of the last `span` steps (the device counter holds
// the number of steps done): one thread per (DFT entry, frequency).
template <typename real>
__global__ __launch_bounds__(256) void k_batch_dft_fold(BatchArgs<real> a, int span, int n_freq) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)a.n_dft * n_freq) return;
  const int64_t s_hi = *a.step;
  dft_fold(a, (int)(t / n_freq), (int)(t % n_freq), s_hi - span, s_hi);
}

__global__ void k_batch_sync_reset(unsigned* bar, int* fail) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    bar[0] = 0;
    *fail = 0;
  }
}

// Persistent kernel: steps [step0, step1) in one launch; per step the march
// over all chunks, a grid barrier, the post pass, a grid barrier; every
// BATCH_DFT_RING steps and at the end the DFT fold and a grid barrier. Every
// wavefront loops over its share of the chunks, every thread over its share
// of the post nodes and fold items. The device step counter is not used.
template <typename real, bool EXACT>
__global__ __launch_bounds__(512) void k_batch_persistent(BatchArgs<real> a, int64_t step0,
                                                          int64_t step1, int n_freq) {
  const int wave = wave_id();
  const int waves = (int)(gridDim.x * blockDim.x) >> 6;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int threads = gridDim.x * blockDim.x;
  int64_t folded = step0;
  for (int64_t st = step0; st < step1; ++st) {
    for (int ci = wave; ci < a.n_chunks; ci += waves) {
      const MarchChunk c = a.chunks[ci];
      const SimDev<real>& s = a.sims[c.sim];
      const bool hsrc = s.src_h.n_groups > 0;
      if (s.pol == 0) {
        if (hsrc) march_chunk<real, EXACT, 0, true>(a, c, st);
        else march_chunk<real, EXACT, 0, false>(a, c, st);
      } else {
        if (hsrc) march_chunk<real, EXACT, 1, true>(a, c, st);
        else march_chunk<real, EXACT, 1, false>(a, c, st);
      }
    }
    if (!grid_barrier(a.bar, a.fail, a.wait_ticks)) return;
    for (int idx = tid; idx < a.n_post; idx += threads) post_node(a, idx, st);
    if (!grid_barrier(a.bar, a.fail, a.wait_ticks)) return;
    if (st + 1 - folded == BATCH_DFT_RING || st + 1 == step1) {
      for (int64_t t = tid; t < (int64_t)a.n_dft * n_freq; t += threads)
        dft_fold(a, (int)(t / n_freq), (int)(t % n_freq), folded, st + 1);
      folded = st + 1;
      if (!grid_barrier(a.bar, a.fail, a.wait_ticks)) return;
    }
  }
}

template <typename real>
void launch_batch_march(const BatchArgs<real>& a, const ChunkGroup& g, bool exact,
                        int waves_per_block, hipStream_t s) {
  if (g.count <= 0) return;
  const int blocks = (g.count + waves_per_block - 1) / waves_per_block;
  const int threads = waves_per_block * 64;
#define PF_MARCH(E, P, H) \
  k_batch_march<real, E, P, H><<<blocks, threads, 0, s>>>(a, g.first, g.count)
  if (exact) {
    if (g.pol == 0) { if (g.hsrc) PF_MARCH(true, 0, true); else PF_MARCH(true, 0, false); }
    else { if (g.hsrc) PF_MARCH(true, 1, true); else PF_MARCH(true, 1, false); }
  } else {
    if (g.pol == 0) { if (g.hsrc) PF_MARCH(false, 0, true); else PF_MARCH(false, 0, false); }
    else { if (g.hsrc) PF_MARCH(false, 1, true); else PF_MARCH(false, 1, false); }
  }
#undef PF_MARCH
}

template <typename real>
void launch_batch_post(const BatchArgs<real>& a, hipStream_t s) {
  k_batch_post<real><<<std::max(1, (a.n_post + 255) / 256), 256, 0, s>>>(a);
}

template <typename real>
void launch_batch_dft_fold(const BatchArgs<real>& a, int span, int n_freq, hipStream_t s) {
  const int64_t n = (int64_t)a.n_dft * n_freq;
  if (n > 0 && span > 0)
    k_batch_dft_fold<real><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, span, n_freq);
}

void launch_batch_sync_reset(unsigned* bar, int* fail, hipStream_t s) {
  k_batch_sync_reset<<<1, 64, 0, s>>>(bar, fail);
}

template <typename real>
hipError_t batch_persistent_capacity(bool exact, int waves_per_block, int* per_cu, int* n_cu) {
  const void* fn = exact ? reinterpret_cast<const void*>(&k_batch_persistent<real, true>)
                         : reinterpret_cast<const void*>(&k_batch_persistent<real, false>);
  int dev = 0;
  *per_cu = 0;
  *n_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fn, waves_per_block * 64, 0);
  if (e != hipSuccess) return e;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  return hipDeviceGetAttribute(n_cu, hipDeviceAttributeMultiprocessorCount, dev);
}

template <typename real>
void launch_batch_persistent(const BatchArgs<real>& a, bool exact, int waves_per_block, int blocks,
                             int64_t step0, int64_t step1, int n_freq, hipStream_t s) {
  if (exact)
    k_batch_persistent<real, true><<<blocks, waves_per_block * 64, 0, s>>>(a, step0, step1, n_freq);
  else
    k_batch_persistent<real, false><<<blocks, waves_per_block * 64, 0, s>>>(a, step0, step1, n_freq);
}

template void launch_batch_march<float>(const BatchArgs<float>&, const ChunkGroup&, bool, int,
                                        hipStream_t);
template void launch_batch_march<double>(const BatchArgs<double>&, const ChunkGroup&, bool, int,
                                         hipStream_t);
template void launch_batch_post<float>(const BatchArgs<float>&, hipStream_t);
template void launch_batch_post<double>(const BatchArgs<double>&, hipStream_t);
template void launch_batch_dft_fold<float>(const BatchArgs<float>&, int, int, hipStream_t);
template void launch_batch_dft_fold<double>(const BatchArgs<double>&, int, int, hipStream_t);
template void launch_batch_persistent<float>(const BatchArgs<float>&, bool, int, int, int64_t,
                                             int64_t, int, hipStream_t);
template void launch_batch_persistent<double>(const BatchArgs<double>&, bool, int, int, int64_t,
                                              int64_t, int, hipStream_t);
template hipError_t batch_persistent_capacity<float>(bool, int, int*, int*);
template hipError_t batch_persistent_capacity<double>(bool, int, int*, int*);

}  // namespace pfhip
