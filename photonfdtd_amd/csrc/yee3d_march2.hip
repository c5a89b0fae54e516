// Two-step march (temporal blocking): one pass advances the fields n -> n+2.
//
// The tile layout is the single-step march's with a two-node apron: wave w
// handles row j0-2+w and lane l z node k0-2+l. Marching over x, iteration p
// runs step A (n -> n+1) at plane p and step B (n+1 -> n+2) at plane p-1:
//
//   A:  H^{n+1/2}(p) from E^n(p), E^n(p+1);  E^{n+1}(p) from H^{n+1/2}(p), (p-1)
//   B:  H^{n+3/2}(p-1) from E^{n+1}(p-1), E^{n+1}(p);  E^{n+2}(p-1) from H^{n+3/2}(p-1), (p-2)
//
// y neighbours go through LDS (rows +1 for the H updates, -1 for the E
// updates) and z neighbours through cross-lane shuffles, so every update is
// valid one node less far out than its inputs: step A's H on rows/lanes
// 0..jn+2, its E on 1..jn+2, step B's H on 1..jn+1, its E (the owned nodes,
// the only ones stored) on 2..jn+1. Level n+1 never leaves registers/LDS.
//
// Everything a recomputed halo node reads must be the old state, so the
// fields, both psi sides and the ADE polarisation are double-buffered: read
// level n, write level n+2. The ADE correction and the E-side soft sources of
// both steps run in the kernel (media must fill boxes; sources are found by
// binary search on the sorted groups); DFT samples of step n are written
// straight into the monitor's ring slot. Each node performs the same
// operations in the same order as two single steps, so the result is bit for
// bit that of the single-step path.
//
// Memory access: every field, ce and psi load and store is a buffer access
// whose masked lanes carry an out-of-range offset (no branch), and the first
// two and last two planes of a chunk run in a checked iteration, so the
// steady loop issues the same memory operations on every path and the next
// plane's loads stay in flight across a whole iteration.
#include <type_traits>

#include "pfhip_internal.h"
#include "yee_common.h"

namespace pfhip {

template <typename real, bool EXACT, int NW, bool PML>
__global__ __launch_bounds__(NW * 64) void k_yee_march2(MarchArgs2<real> a) {
  constexpr uint32_t B = sizeof(real);
  constexpr int XS = MARCH_MAX_XCHUNK + 3;
  __shared__ real sEx[NW * 64], sEz[NW * 64], sHx[NW * 64], sHz[NW * 64];
  __shared__ real sX[4][XS];    // b_e, c_e, b_h, c_h along x for the chunk
  // psi of step A carried to step B, one private slot per thread: H side
  // (HxY HxZ HyZ HyX HzX HzY) then E side (ExY ExZ EyZ EyX EzX EzY)
  __shared__ real sPsi[PML ? 12 : 1][NW * 64];
  __shared__ int sAde[MARCH2_MAX_ADE][2];    // x range [lo, hi) of each medium
  __shared__ int sDft[MARCH2_MAX_DFT][3];    // x start, stride, count of each monitor
  if ((int)blockIdx.x >= a.ntiles) return;
  const MarchTile tile = reinterpret_cast<const MarchTile*>(a.tiles)[blockIdx.x];
  if (tile.i1 <= tile.i0) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nyz = (size_t)ny * nz;
  // Field and ce arrays use rows of a.zrow slots (padded rows when a.zpad:
  // node k in slot zslot(k), each 60-node z tile (a.zpad = 60) owning one aligned 64-slot
  // block); psi arrays keep dense rows of nz.
  const size_t fplane = (size_t)ny * a.zrow, fcell = (size_t)nx * fplane;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int jn = tile.jn, kn = tile.kn;
  const int j = tile.j0 - 2 + w, k = tile.k0 - 2 + l;
  const bool inb = j >= 0 && j < ny && k >= 0 && k < nz;
  const bool lrole = inb && w <= jn + 3 && l <= kn + 3;   // holds E^n
  const bool hA = lrole && w <= jn + 2 && l <= kn + 2;     // step-A H
  const bool eA = hA && w >= 1 && l >= 1;                  // step-A E
  const bool hB = eA && w <= jn + 1 && l <= kn + 1;        // step-B H
  const bool own = hB && w >= 2 && l >= 2;                 // step-B E; stored
  const int jc = min(max(j, 0), ny - 1), kc = min(max(k, 0), nz - 1);
  const int i0 = tile.i0, i1 = tile.i1;
  const int ib = i0 > 2 ? i0 - 2 : 0;           // first plane of step A
  const int ie = min(i1, nx - 1);               // last plane of step A
  const int qb = i0 > 1 ? i0 - 1 : 0;           // first plane of step B
  const int64_t n = a.step;

  const real* Ein0 = a.Lin;
  const real* Ein1 = a.Lin + fcell;
  const real* Ein2 = a.Lin + 2 * fcell;
  const real* Hin0 = a.Lin + 3 * fcell;
  const real* Hin1 = a.Lin + 4 * fcell;
  const real* Hin2 = a.Lin + 5 * fcell;
  real* Eo0 = a.Lout;
  real* Eo1 = a.Lout + fcell;
  real* Eo2 = a.Lout + 2 * fcell;
  real* Ho0 = a.Lout + 3 * fcell;
  real* Ho1 = a.Lout + 4 * fcell;
  real* Ho2 = a.Lout + 5 * fcell;
  const size_t sy_e = (size_t)nx * a.pe[1] * nz, sz_e = (size_t)nx * ny * a.pe[2],
               sx_e = (size_t)a.pe[0] * nyz;
  const size_t sy_h = (size_t)nx * a.ph[1] * nz, sz_h = (size_t)nx * ny * a.ph[2],
               sx_h = (size_t)a.ph[0] * nyz;
  // slot offsets: E: ExY ExZ EyZ EyX EzX EzY ; H: HxY HxZ HyZ HyX HzX HzY
  const size_t eoff[6] = {0, sy_e, sy_e + sz_e, sy_e + 2 * sz_e, sy_e + 2 * sz_e + sx_e,
                          sy_e + 2 * sz_e + 2 * sx_e};
  const size_t hoff[6] = {0, sy_h, sy_h + sz_h, sy_h + 2 * sz_h, sy_h + 2 * sz_h + sx_h,
                          sy_h + 2 * sz_h + 2 * sx_h};
  const real* coef_y = a.coef + 4 * (size_t)nx;
  const real* coef_z = a.coef + 4 * (size_t)(nx + ny);

  for (int t = tid; PML && t < 4 * (ie - ib + 1); t += NW * 64) {
    const int q = t / (ie - ib + 1), ii = ib + t % (ie - ib + 1);
    sX[q][ii - ib] = a.coef[(size_t)q * nx + ii];
  }

  // media and monitors whose (y, z) extent holds this lane's column; their
  // x ranges go to LDS, so the plane loop tests them without memory loads
  uint32_t ade_mask = 0, dft_mask = 0;
  for (int m = 0; m < a.n_ade; ++m) {
    const auto* M = gptr(a.ade) + m;
    const int lo1 = M->lo[1], hi1 = M->hi[1], lo2 = M->lo[2], hi2 = M->hi[2];
    if (M->n_poles > 0 && j >= lo1 && j < hi1 && k >= lo2 && k < hi2) ade_mask |= 1u << m;
    if (tid == 0) {
      sAde[m][0] = M->lo[0];
      sAde[m][1] = M->hi[0];
    }
  }
  for (int m = 0; m < a.n_dft; ++m) {
    const auto* D = gptr(a.dft) + m;
    const int dj = j - D->start[1], dk = k - D->start[2];
    const int sj = D->stride[1], sk = D->stride[2];
    if (dj >= 0 && dk >= 0 && dj % sj == 0 && dk % sk == 0 && dj / sj < D->count[1] &&
        dk / sk < D->count[2])
      dft_mask |= 1u << m;
    if (tid == 0) {
      sDft[m][0] = D->start[0];
      sDft[m][1] = D->stride[0];
      sDft[m][2] = D->count[0];
    }
  }

  const Slabs<real> ye{a.slab_e[1][0], a.slab_e[1][1]}, yh{a.slab_h[1][0], a.slab_h[1][1]};
  const Slabs<real> ze{a.slab_e[2][0], a.slab_e[2][1]}, zh{a.slab_h[2][0], a.slab_h[2][1]};
  const Slabs<real> xe{a.slab_e[0][0], a.slab_e[0][1]}, xh{a.slab_h[0][0], a.slab_h[0][1]};
  const bool pyh = PML && hA && yh.in(jc), pzh = PML && hA && zh.in(kc);
  const bool pye = PML && eA && ye.in(jc), pze = PML && eA && ze.in(kc);
  const real bye = pye ? coef_y[jc] : real(0), cye = pye ? coef_y[ny + jc] : real(0);
  const real byh = pyh ? coef_y[2 * ny + jc] : real(0), cyh = pyh ? coef_y[3 * ny + jc] : real(0);
  const real bze = pze ? coef_z[kc] : real(0), cze = pze ? coef_z[nz + kc] : real(0);
  const real bzh = pzh ? coef_z[2 * nz + kc] : real(0), czh = pzh ? coef_z[3 * nz + kc] : real(0);
  const uint32_t col = (uint32_t)(jc * nz + kc) * B;
  const uint32_t cyh_o = (uint32_t)(yh.m(jc) * nz + kc) * B;
  const uint32_t czh_o = (uint32_t)(jc * a.ph[2] + zh.m(kc)) * B;
  const uint32_t cye_o = (uint32_t)(ye.m(jc) * nz + kc) * B;
  const uint32_t cze_o = (uint32_t)(jc * a.pe[2] + ze.m(kc)) * B;
  const size_t ys_h = (size_t)a.ph[1] * nz, zs_h = (size_t)ny * a.ph[2];
  const size_t ys_e = (size_t)a.pe[1] * nz, zs_e = (size_t)ny * a.pe[2];
  // lane offsets per role (BUF_NONE where the role does not hold); o_*
  // address psi (dense rows), f_* the fields and ce
  const uint32_t o_h = hA ? col : BUF_NONE;
  const uint32_t o_e = eA ? col : BUF_NONE, o_own = own ? col : BUF_NONE;
  const uint32_t colf = (uint32_t)(jc * a.zrow + zslot(kc, a.zpad)) * B;
  const uint32_t f_l = lrole ? colf : BUF_NONE, f_h = hA ? colf : BUF_NONE;
  const uint32_t f_e = eA ? colf : BUF_NONE;
  // field stores: the owned nodes; with padded rows every lane of an owned
  // row stores into the tile's block (lanes 0, 1, 62, 63 and lanes past the
  // grid's end into its pad slots), so a row store covers whole lines
  const bool orow = j >= 0 && j < ny && w >= 2 && w <= jn + 1;
  const uint32_t f_own =
      a.zpad ? (orow ? (uint32_t)(jc * a.zrow + tile.k0 / a.zpad * ZPAD_BLOCK +
                                  (l < 2 ? a.zpad + l : l <= a.zpad + 1 ? l - 2 : l)) * B
                     : BUF_NONE)
             : (own ? colf : BUF_NONE);
  const uint32_t o_pyh = pyh ? cyh_o : BUF_NONE, o_pzh = pzh ? czh_o : BUF_NONE;
  const uint32_t o_pye = pye ? cye_o : BUF_NONE, o_pze = pze ? cze_o : BUF_NONE;
  // psi stores of step B: owned nodes only
  const uint32_t s_pyh = own ? o_pyh : BUF_NONE, s_pzh = own ? o_pzh : BUF_NONE;
  const uint32_t s_pye = own ? o_pye : BUF_NONE, s_pze = own ? o_pze : BUF_NONE;

  struct Psi {
    real ph[6], pe[6];
  };
  struct Plane {
    real e[3];    // E^n(i+1)
    real h[3];    // H^{n-1/2}(i)
    real ce;
    Psi q;
  };

  auto fetch = [&](Plane& P, int i) __attribute__((always_inline)) {
    const size_t o = (size_t)i * fplane;
    const size_t on = (size_t)min(i + 1, nx - 1) * fplane;
    const uint32_t oe = i + 1 < nx ? f_l : BUF_NONE;
    P.e[0] = bld(Ein0 + on, oe);
    P.e[1] = bld(Ein1 + on, oe);
    P.e[2] = bld(Ein2 + on, oe);
    P.h[0] = bld(Hin0 + o, f_h);
    P.h[1] = bld(Hin1 + o, f_h);
    P.h[2] = bld(Hin2 + o, f_h);
    P.ce = bld(a.ce + o, f_e);
    Psi& Q = P.q;
    if (!PML) {
      for (int q = 0; q < 6; ++q) Q.ph[q] = Q.pe[q] = real(0);
      return;
    }
    Q.ph[0] = bld(a.psiHin + hoff[0] + i * ys_h, o_pyh);
    Q.ph[5] = bld(a.psiHin + hoff[5] + i * ys_h, o_pyh);
    Q.ph[1] = bld(a.psiHin + hoff[1] + i * zs_h, o_pzh);
    Q.ph[2] = bld(a.psiHin + hoff[2] + i * zs_h, o_pzh);
    const bool xhi = xh.in(i), xei = xe.in(i);
    const size_t xoh = (size_t)(xhi ? xh.m(i) : 0) * nyz, xoe = (size_t)(xei ? xe.m(i) : 0) * nyz;
    Q.ph[3] = bld(a.psiHin + hoff[3] + xoh, xhi ? o_h : BUF_NONE);
    Q.ph[4] = bld(a.psiHin + hoff[4] + xoh, xhi ? o_h : BUF_NONE);
    Q.pe[0] = bld(a.psiEin + eoff[0] + i * ys_e, o_pye);
    Q.pe[5] = bld(a.psiEin + eoff[5] + i * ys_e, o_pye);
    Q.pe[1] = bld(a.psiEin + eoff[1] + i * zs_e, o_pze);
    Q.pe[2] = bld(a.psiEin + eoff[2] + i * zs_e, o_pze);
    Q.pe[3] = bld(a.psiEin + eoff[3] + xoe, xei ? o_e : BUF_NONE);
    Q.pe[4] = bld(a.psiEin + eoff[4] + xoe, xei ? o_e : BUF_NONE);
  };

  // E-side soft sources of component c at a cell for time index t
  // (k_inject's arithmetic: add in float64, round per dipole).
  auto add_src = [&](int c, int64_t cell, int64_t t, real v) __attribute__((always_inline)) -> real {
    const int lo = c == 0 ? a.src_beg[0] : c == 1 ? a.src_beg[1] : a.src_beg[2];
    const int hi = c == 0 ? a.src_beg[1] : c == 1 ? a.src_beg[2] : a.src_beg[3];
    int L = lo, H = hi;
    while (L < H) {
      const int mid = (L + H) >> 1;
      if (gptr(a.src_cell)[mid] < cell) L = mid + 1;
      else H = mid;
    }
    if (L >= hi || gptr(a.src_cell)[L] != cell) return v;
    const auto* beg = gptr(a.src_begin);
    for (int s = beg[L]; s < beg[L + 1]; ++s)
      v = (real)((double)v + gptr(a.src_amp)[s] *
                                 gptr(a.src_wave)[(int64_t)gptr(a.src_row)[s] * a.n_steps + t]);
    return v;
  };

  // ADE of step A: P^{n+1} from level n; the owner keeps it in Pprev_out.
  auto ade_a = [&](int p, const real* en, real* e, bool keep) __attribute__((always_inline)) {
    for (int m = 0; m < a.n_ade; ++m) {
      if (!((ade_mask >> m) & 1u) || p < sAde[m][0] || p >= sAde[m][1]) continue;
      const auto* M = gptr(a.ade) + m;
      const int lo1 = M->lo[1], hi1 = M->hi[1], lo2 = M->lo[2], hi2 = M->hi[2];
      const int64_t t = ((int64_t)(p - sAde[m][0]) * (hi1 - lo1) + (j - lo1)) * (hi2 - lo2) + (k - lo2);
      const int n_poles = M->n_poles;
      const int64_t nn = M->n;
      const auto* abc = gptr(M->abc);
      const auto* Pc = gptr(M->Pcur_in);
      const auto* Pp = gptr(M->Pprev_in);
      auto* Po = gptr(M->Pprev_out);
      const real inv_eps = M->inv_eps;
      for (int c = 0; c < 3; ++c) {
        real delta = 0;
        for (int q = 0; q < n_poles; ++q) {
          const int64_t off = ((int64_t)q * 3 + c) * nn + t;
          const real pc = Pc[off], pv = Pp[off];
          const real pn = abc[3 * q] * pc + abc[3 * q + 1] * pv + abc[3 * q + 2] * en[c];
          const real inc = pn - pc;
          delta = q == 0 ? inc : delta + inc;
          if (keep) Po[off] = pn;
        }
        e[c] = e[c] - inv_eps * delta;
      }
    }
  };
  // ADE of step B (owned nodes): P^{n+2} from P^{n+1} (Pprev_out) and P^n.
  auto ade_b = [&](int q, const real* en1, real* e) __attribute__((always_inline)) {
    for (int m = 0; m < a.n_ade; ++m) {
      if (!((ade_mask >> m) & 1u) || q < sAde[m][0] || q >= sAde[m][1]) continue;
      const auto* M = gptr(a.ade) + m;
      const int lo1 = M->lo[1], hi1 = M->hi[1], lo2 = M->lo[2], hi2 = M->hi[2];
      const int64_t t = ((int64_t)(q - sAde[m][0]) * (hi1 - lo1) + (j - lo1)) * (hi2 - lo2) + (k - lo2);
      const int n_poles = M->n_poles;
      const int64_t nn = M->n;
      const auto* abc = gptr(M->abc);
      const auto* Pc = gptr(M->Pcur_in);
      const auto* P1 = gptr(M->Pprev_out);
      auto* Po = gptr(M->Pcur_out);
      const real inv_eps = M->inv_eps;
      for (int c = 0; c < 3; ++c) {
        real delta = 0;
        for (int r = 0; r < n_poles; ++r) {
          const int64_t off = ((int64_t)r * 3 + c) * nn + t;
          const real pc = P1[off], pv = Pc[off];
          const real pn = abc[3 * r] * pc + abc[3 * r + 1] * pv + abc[3 * r + 2] * en1[c];
          const real inc = pn - pc;
          delta = r == 0 ? inc : delta + inc;
          Po[off] = pn;
        }
        e[c] = e[c] - inv_eps * delta;
      }
    }
  };

  // DFT samples of step n at an owned node of an owned plane.
  auto capture = [&](int p, real ex, real ey, real ez, real hx, real hy, real hz) __attribute__((always_inline)) {
    for (int m = 0; m < a.n_dft; ++m) {
      const uint32_t sl = (a.dft_slots >> (8 * m)) & 0xFFu;
      if (sl == 0xFFu || !((dft_mask >> m) & 1u)) continue;
      const int di = p - sDft[m][0], si = sDft[m][1];
      if (di < 0 || di % si != 0 || di / si >= sDft[m][2]) continue;
      const auto* D = gptr(a.dft) + m;
      const int c1 = D->count[1], c2 = D->count[2];
      const size_t nbox = (size_t)sDft[m][2] * c1 * c2;
      const size_t r = ((size_t)(di / si) * c1 + (j - D->start[1]) / D->stride[1]) * c2 +
                       (k - D->start[2]) / D->stride[2];
      auto* slot = gptr(D->ring + (int64_t)sl * D->slot_elems);
      const int nc = D->n_comp;
      for (int c = 0; c < nc; ++c) {
        const int comp = D->comps[c];
        const real v = comp == 0 ? ex : comp == 1 ? ey : comp == 2 ? ez
                     : comp == 3 ? hx : comp == 4 ? hy : hz;
        slot[c * nbox + r] = v;
      }
    }
  };

  real ecur[3];                    // E^n(p)
  {
    const size_t o = (size_t)ib * fplane;
    ecur[0] = bld(Ein0 + o, f_l);
    ecur[1] = bld(Ein1 + o, f_l);
    ecur[2] = bld(Ein2 + o, f_l);
  }
  real hA_y = 0, hA_z = 0;         // H^{n+1/2}(p-1): Hy, Hz
  real hB_y = 0, hB_z = 0;         // H^{n+3/2}(q-1): Hy, Hz
  real cE[3] = {0, 0, 0}, cH[3] = {0, 0, 0}, cCe = 0;   // step-A results of plane p-1
  const real K_h = -a.ch;

  // Step A at plane p: results in nE, nH, nPh, nPe.
  auto step_a = [&](int p, const Plane& P, real* nE, real* nH, real* nPh, real* nPe) __attribute__((always_inline)) {
    const Psi& Q = P.q;
    sEx[tid] = ecur[0];
    sEz[tid] = ecur[2];
    lds_barrier();
    const real ex_jp = w < NW - 1 ? sEx[min(tid + 64, NW * 64 - 1)] : real(0);
    const real ez_jp = w < NW - 1 ? sEz[min(tid + 64, NW * 64 - 1)] : real(0);
    const real ex_kp = __shfl_down(ecur[0], 1);
    const real ey_kp = __shfl_down(ecur[1], 1);
    const bool xh_in = p < nx - 1, yh_in = j < ny - 1, zh_in = k < nz - 1;
    const bool pxh = PML && xh.in(p);
    const real bxh = PML ? sX[2][p - ib] : real(0), cxh = PML ? sX[3][p - ib] : real(0);
    real hx = P.h[0], hy = P.h[1], hz = P.h[2];
    for (int q = 0; q < 6; ++q) nPh[q] = Q.ph[q];
    if (hA) {
      if (yh_in && zh_in) {
        const real dA = diff<real, EXACT>(ez_jp, ecur[2], a.d[1], a.inv_d[1]);
        const real dB = diff<real, EXACT>(ey_kp, ecur[1], a.d[2], a.inv_d[2]);
        hx = curl_update(hx, K_h, true, dA, pyh, Q.ph[0], byh, cyh, &nPh[0], true, dB, pzh,
                         Q.ph[1], bzh, czh, &nPh[1]);
      }
      if (zh_in && xh_in) {
        const real dA = diff<real, EXACT>(ex_kp, ecur[0], a.d[2], a.inv_d[2]);
        const real dB = diff<real, EXACT>(P.e[2], ecur[2], a.d[0], a.inv_d[0]);
        hy = curl_update(hy, K_h, true, dA, pzh, Q.ph[2], bzh, czh, &nPh[2], true, dB, pxh,
                         Q.ph[3], bxh, cxh, &nPh[3]);
      }
      if (xh_in && yh_in) {
        const real dA = diff<real, EXACT>(P.e[1], ecur[1], a.d[0], a.inv_d[0]);
        const real dB = diff<real, EXACT>(ex_jp, ecur[0], a.d[1], a.inv_d[1]);
        hz = curl_update(hz, K_h, true, dA, pxh, Q.ph[4], bxh, cxh, &nPh[4], true, dB, pyh,
                         Q.ph[5], byh, cyh, &nPh[5]);
      }
    }
    sHx[tid] = hx;
    sHz[tid] = hz;
    lds_barrier();
    const real hx_jm = w > 0 ? sHx[tid - 64] : real(0);
    const real hz_jm = w > 0 ? sHz[tid - 64] : real(0);
    const real hx_km = __shfl_up(hx, 1);
    const real hy_km = __shfl_up(hy, 1);
    real e[3] = {ecur[0], ecur[1], ecur[2]};
    for (int q = 0; q < 6; ++q) nPe[q] = Q.pe[q];
    if (eA) {
      const bool xe_in = p >= 1, ye_in = j >= 1, ze_in = k >= 1;
      const bool pxe = PML && xe.in(p);
      const real bxe = PML ? sX[0][p - ib] : real(0), cxe = PML ? sX[1][p - ib] : real(0);
      const real K = P.ce;
      if (ye_in && ze_in) {
        const real dA = diff<real, EXACT>(hz, hz_jm, a.d[1], a.inv_d[1]);
        const real dB = diff<real, EXACT>(hy, hy_km, a.d[2], a.inv_d[2]);
        e[0] = curl_update(e[0], K, true, dA, pye, Q.pe[0], bye, cye, &nPe[0], true, dB, pze,
                           Q.pe[1], bze, cze, &nPe[1]);
      }
      if (ze_in && xe_in) {
        const real dA = diff<real, EXACT>(hx, hx_km, a.d[2], a.inv_d[2]);
        const real dB = diff<real, EXACT>(hz, hA_z, a.d[0], a.inv_d[0]);
        e[1] = curl_update(e[1], K, true, dA, pze, Q.pe[2], bze, cze, &nPe[2], true, dB, pxe,
                           Q.pe[3], bxe, cxe, &nPe[3]);
      }
      if (xe_in && ye_in) {
        const real dA = diff<real, EXACT>(hy, hA_y, a.d[0], a.inv_d[0]);
        const real dB = diff<real, EXACT>(hx, hx_jm, a.d[1], a.inv_d[1]);
        e[2] = curl_update(e[2], K, true, dA, pxe, Q.pe[4], bxe, cxe, &nPe[4], true, dB, pye,
                           Q.pe[5], bye, cye, &nPe[5]);
      }
      const bool owned_plane = p >= i0 && p < i1;
      if (ade_mask) ade_a(p, ecur, e, own && owned_plane);
      if (p >= a.src_i_lo && p <= a.src_i_hi) {
        const int64_t cell = (int64_t)p * (int64_t)fplane + (int64_t)j * a.zrow + zslot(k, a.zpad);
        for (int c = 0; c < 3; ++c) e[c] = add_src(c, cell, n, e[c]);
      }
      if (dft_mask && own && owned_plane) capture(p, e[0], e[1], e[2], hx, hy, hz);
    }
    hA_y = hy;
    hA_z = hz;
    for (int c = 0; c < 3; ++c) nE[c] = e[c];
    nH[0] = hx;
    nH[1] = hy;
    nH[2] = hz;
  };

  // Step B at plane q from the carried step-A results of q and E^{n+1}(q+1).
  auto step_b = [&](int q, const real* eNext) __attribute__((always_inline)) {
    sEx[tid] = cE[0];
    sEz[tid] = cE[2];
    lds_barrier();
    const real ex_jp = w < NW - 1 ? sEx[min(tid + 64, NW * 64 - 1)] : real(0);
    const real ez_jp = w < NW - 1 ? sEz[min(tid + 64, NW * 64 - 1)] : real(0);
    const real ex_kp = __shfl_down(cE[0], 1);
    const real ey_kp = __shfl_down(cE[1], 1);
    const bool xh_in = q < nx - 1, yh_in = j < ny - 1, zh_in = k < nz - 1;
    const bool pxh = PML && xh.in(q);
    const real bxh = PML ? sX[2][q - ib] : real(0), cxh = PML ? sX[3][q - ib] : real(0);
    real hx = cH[0], hy = cH[1], hz = cH[2];
    real nps[6] = {0, 0, 0, 0, 0, 0};
    real cPh[6] = {0, 0, 0, 0, 0, 0}, cPe[6] = {0, 0, 0, 0, 0, 0};
    if (PML)
      for (int q6 = 0; q6 < 6; ++q6) {
        cPh[q6] = sPsi[q6][tid];
        cPe[q6] = sPsi[6 + q6][tid];
      }
    if (hB) {
      if (yh_in && zh_in) {
        const real dA = diff<real, EXACT>(ez_jp, cE[2], a.d[1], a.inv_d[1]);
        const real dB = diff<real, EXACT>(ey_kp, cE[1], a.d[2], a.inv_d[2]);
        hx = curl_update(hx, K_h, true, dA, pyh, cPh[0], byh, cyh, &nps[0], true, dB, pzh,
                         cPh[1], bzh, czh, &nps[1]);
      }
      if (zh_in && xh_in) {
        const real dA = diff<real, EXACT>(ex_kp, cE[0], a.d[2], a.inv_d[2]);
        const real dB = diff<real, EXACT>(eNext[2], cE[2], a.d[0], a.inv_d[0]);
        hy = curl_update(hy, K_h, true, dA, pzh, cPh[2], bzh, czh, &nps[2], true, dB, pxh,
                         cPh[3], bxh, cxh, &nps[3]);
      }
      if (xh_in && yh_in) {
        const real dA = diff<real, EXACT>(eNext[1], cE[1], a.d[0], a.inv_d[0]);
        const real dB = diff<real, EXACT>(ex_jp, cE[0], a.d[1], a.inv_d[1]);
        hz = curl_update(hz, K_h, true, dA, pxh, cPh[4], bxh, cxh, &nps[4], true, dB, pyh,
                         cPh[5], byh, cyh, &nps[5]);
      }
    }
    const bool oq = q >= i0 && q < i1;           // owned plane (uniform)
    const bool store = own && oq;
    const uint32_t os = oq ? o_own : BUF_NONE;       // psi (dense rows)
    const uint32_t fos = oq ? f_own : BUF_NONE;      // fields
    {
      const size_t o = (size_t)q * fplane;
      bst(Ho0 + o, fos, hx);
      bst(Ho1 + o, fos, hy);
      bst(Ho2 + o, fos, hz);
      if (PML) {
        const bool yz = yh_in && zh_in, zx = zh_in && xh_in, xy = xh_in && yh_in;
        const bool xq = xh.in(q);
        const size_t xo = (size_t)(xq ? xh.m(q) : 0) * nyz;
        bst(a.psiHout + hoff[0] + q * ys_h, yz && oq ? s_pyh : BUF_NONE, nps[0]);
        bst(a.psiHout + hoff[1] + q * zs_h, yz && oq ? s_pzh : BUF_NONE, nps[1]);
        bst(a.psiHout + hoff[2] + q * zs_h, zx && oq ? s_pzh : BUF_NONE, nps[2]);
        bst(a.psiHout + hoff[3] + xo, xq && zx ? os : BUF_NONE, nps[3]);
        bst(a.psiHout + hoff[4] + xo, xq && xy ? os : BUF_NONE, nps[4]);
        bst(a.psiHout + hoff[5] + q * ys_h, xy && oq ? s_pyh : BUF_NONE, nps[5]);
      }
    }
    sHx[tid] = hx;
    sHz[tid] = hz;
    lds_barrier();
    const real hx_jm = w > 0 ? sHx[tid - 64] : real(0);
    const real hz_jm = w > 0 ? sHz[tid - 64] : real(0);
    const real hx_km = __shfl_up(hx, 1);
    const real hy_km = __shfl_up(hy, 1);
    const bool xe_in = q >= 1, ye_in = j >= 1, ze_in = k >= 1;
    const bool pxe = PML && xe.in(q);
    real e[3] = {cE[0], cE[1], cE[2]};
    real npe[6] = {0, 0, 0, 0, 0, 0};
    if (own) {
      const real bxe = PML ? sX[0][q - ib] : real(0), cxe = PML ? sX[1][q - ib] : real(0);
      const real K = cCe;
      if (ye_in && ze_in) {
        const real dA = diff<real, EXACT>(hz, hz_jm, a.d[1], a.inv_d[1]);
        const real dB = diff<real, EXACT>(hy, hy_km, a.d[2], a.inv_d[2]);
        e[0] = curl_update(e[0], K, true, dA, pye, cPe[0], bye, cye, &npe[0], true, dB, pze,
                           cPe[1], bze, cze, &npe[1]);
      }
      if (ze_in && xe_in) {
        const real dA = diff<real, EXACT>(hx, hx_km, a.d[2], a.inv_d[2]);
        const real dB = diff<real, EXACT>(hz, hB_z, a.d[0], a.inv_d[0]);
        e[1] = curl_update(e[1], K, true, dA, pze, cPe[2], bze, cze, &npe[2], true, dB, pxe,
                           cPe[3], bxe, cxe, &npe[3]);
      }
      if (xe_in && ye_in) {
        const real dA = diff<real, EXACT>(hy, hB_y, a.d[0], a.inv_d[0]);
        const real dB = diff<real, EXACT>(hx, hx_jm, a.d[1], a.inv_d[1]);
        e[2] = curl_update(e[2], K, true, dA, pxe, cPe[4], bxe, cxe, &npe[4], true, dB, pye,
                           cPe[5], bye, cye, &npe[5]);
      }
    }
    if (store && ade_mask) ade_b(q, cE, e);
    if (store && q >= a.src_i_lo && q <= a.src_i_hi) {
      const int64_t cell = (int64_t)q * (int64_t)fplane + (int64_t)j * a.zrow + zslot(k, a.zpad);
      for (int c = 0; c < 3; ++c) e[c] = add_src(c, cell, n + 1, e[c]);
    }
    {
      const size_t o = (size_t)q * fplane;
      bst(Eo0 + o, fos, e[0]);
      bst(Eo1 + o, fos, e[1]);
      bst(Eo2 + o, fos, e[2]);
      if (PML) {
        const bool yz = ye_in && ze_in, zx = ze_in && xe_in, xy = xe_in && ye_in;
        const size_t xo = (size_t)(pxe ? xe.m(q) : 0) * nyz;
        bst(a.psiEout + eoff[0] + q * ys_e, yz && oq ? s_pye : BUF_NONE, npe[0]);
        bst(a.psiEout + eoff[1] + q * zs_e, yz && oq ? s_pze : BUF_NONE, npe[1]);
        bst(a.psiEout + eoff[2] + q * zs_e, zx && oq ? s_pze : BUF_NONE, npe[2]);
        bst(a.psiEout + eoff[3] + xo, pxe && zx ? os : BUF_NONE, npe[3]);
        bst(a.psiEout + eoff[4] + xo, pxe && xy ? os : BUF_NONE, npe[4]);
        bst(a.psiEout + eoff[5] + q * ys_e, xy && oq ? s_pye : BUF_NONE, npe[5]);
      }
    }
    hB_y = hy;
    hB_z = hz;
  };

  // Iteration p: step A at p (p <= ie), then step B at p - 1; the loads of
  // plane p + 1 are issued first. FULL iterations (every plane but the first
  // two and the last two of the chunk) do all three unconditionally.
  auto iter = [&](auto full_t, int p, const Plane& P, Plane& Next) __attribute__((always_inline)) {
    constexpr bool FULL = decltype(full_t)::value;
    const bool doA = FULL || p <= ie;
    const bool doB = FULL || (p - 1 >= qb && p - 1 < i1);
    if (FULL || (doA && p + 1 <= ie)) fetch(Next, p + 1);
    real nE[3] = {0, 0, 0}, nH[3] = {0, 0, 0}, nPh[6], nPe[6];
    if (doA) step_a(p, P, nE, nH, nPh, nPe);
    if (doB) step_b(p - 1, nE);
    if (doA) {
      for (int c = 0; c < 3; ++c) {
        cE[c] = nE[c];
        cH[c] = nH[c];
        ecur[c] = P.e[c];
      }
      if (PML)
        for (int q = 0; q < 6; ++q) {
          sPsi[q][tid] = nPh[q];
          sPsi[6 + q][tid] = nPe[q];
        }
      cCe = P.ce;
    }
  };
  using Checked = std::integral_constant<bool, false>;
  using Full = std::integral_constant<bool, true>;
  Plane A, Bq;
  fetch(A, ib);
  int p = ib;
  for (; p < ib + 2; ++p) {         // planes ib, ib+1 (ie >= ib, so both run)
    iter(Checked{}, p, A, Bq);
    const Plane t = A;
    A = Bq;
    Bq = t;
  }
  // planes ib+2 .. ie-1 in the steady loop, an even number of them; the
  // rest in the checked loop
  const int n_full = max(0, ie - ib - 2) & ~1;
  __builtin_amdgcn_s_waitcnt(0);
  for (int s = 0; s < n_full; s += 2, p += 2) {
    iter(Full{}, p, A, Bq);
    iter(Full{}, p + 1, Bq, A);
  }
  for (; p <= ie + 1; ++p) {
    iter(Checked{}, p, A, Bq);
    const Plane t = A;
    A = Bq;
    Bq = t;
  }
}

template <typename real, int NW>
static void launch_march2_nw(const MarchArgs2<real>& a, bool exact, bool pml, hipStream_t s) {
  if (pml) {
    if (exact) k_yee_march2<real, true, NW, true><<<a.ntiles, NW * 64, 0, s>>>(a);
    else k_yee_march2<real, false, NW, true><<<a.ntiles, NW * 64, 0, s>>>(a);
  } else {
    if (exact) k_yee_march2<real, true, NW, false><<<a.ntiles, NW * 64, 0, s>>>(a);
    else k_yee_march2<real, false, NW, false><<<a.ntiles, NW * 64, 0, s>>>(a);
  }
}

template <typename real>
void launch_march2(const MarchArgs2<real>& a, bool exact, bool pml, int nw, hipStream_t s) {
  if (a.ntiles == 0) return;
  if (nw == 12) launch_march2_nw<real, 12>(a, exact, pml, s);
  else launch_march2_nw<real, 10>(a, exact, pml, s);
}

template void launch_march2<float>(const MarchArgs2<float>&, bool, bool, int, hipStream_t);
template void launch_march2<double>(const MarchArgs2<double>&, bool, bool, int, hipStream_t);

}  // namespace pfhip
