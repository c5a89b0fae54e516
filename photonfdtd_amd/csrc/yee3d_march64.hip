// 3-D march on dense rows with 64 owned z nodes per wavefront.
//
// Layout: every field and ce array is C order with rows of `zrow` slots
// (zrow a multiple of 64, node k in slot k, slots past nz unused), so a
// wavefront's 64 lanes hold 64 consecutive nodes of one row and its row
// loads and stores cover whole cache lines. Tiles start on 64-node
// boundaries and own 64 z nodes (a row's last tile may own fewer; its stores
// are rounded up to 32 lanes, into slots past nz).
//
// A workgroup is NW wavefronts; wave w handles row j0-1+w: wave 0 is the low
// H apron row, wave NW-1 only supplies E(j+1), rows 1..jn are owned. The x
// march, the LDS row exchange (one barrier per plane, double-buffered by
// plane parity), the three rotating plane buffers with two planes of
// prefetch and the per-node arithmetic are those of k_yee_march. What
// differs is the z apron: the H update of the tile's last node needs
// E(k0+kn) and the E update of its first node needs H^{n+1/2}(k0-1). Both
// come from one gather load per plane in which lane q fetches one apron
// value (E at k0+kn and k0-1, H and the H-side psi at k0-1, Ez of the next
// row at k0-1); they are read out with readlane and H^{n+1/2}(k0-1) is
// recomputed from them with the owner's arithmetic (same operands, same
// order, so the same bits as the owning tile).
#include <stdexcept>

#include "pfhip_internal.h"
#include "yee_common.h"

namespace pfhip {

namespace {

template <typename real, bool EXACT>
__device__ __forceinline__ real dd(real hi, real lo, real d, real inv_d) {
  if constexpr (EXACT) return (hi - lo) / d;
  else return (hi - lo) * inv_d;
}

// readlane of a float/double register (lane index wave-uniform).
__device__ __forceinline__ float lane_value(float v, int q) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), q));
}
__device__ __forceinline__ double lane_value(double v, int q) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), q);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), q);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Apron gather slots (lane q of the gather register).
enum Apron {
  A_EX_UP = 0,   // E^n(i+1) at k0+kn: Ex, Ey
  A_EY_UP,
  A_EX_LO,       // E^n(i+1) at k0-1: Ex, Ey, Ez
  A_EY_LO,
  A_EZ_LO,
  A_EZ_LO_JP,    // E^n(i+1), row j+1, at k0-1: Ez
  A_HX_LO,       // H^{n-1/2}(i) at k0-1: Hx, Hy
  A_HY_LO,
  A_PHXY,        // H-side psi (i) at k0-1: HxY HxZ HyZ HyX
  A_PHXZ,
  A_PHYZ,
  A_PHYX,
  A_COUNT
};

}  // namespace

constexpr int M64_XS = MARCH_MAX_XCHUNK + 1;

// LDS of one march workgroup: the row exchange of E and H (double-buffered by
// plane parity) and the x-axis CPML coefficients of the tile's planes.
template <typename real, int NW>
struct M64Lds {
  real ex[2][NW * 64], ez[2][NW * 64], hx[2][NW * 64], hz[2][NW * 64];
  real x[4][M64_XS];
};

// One tile of the march: x planes [i0, i1) of rows j0..j0+jn-1, z nodes
// k0..k0+kn-1. Every wave of the workgroup must call it (it holds LDS
// barriers); it returns with the workgroup's LDS still in use by other
// waves, so a caller that runs another tile next separates the two with a
// barrier.
template <typename real, bool EXACT, int NW, int PMASK, int DIST, bool BATCH>
__device__ __forceinline__ void m64_tile(const MarchArgs<real>& arg, const MarchTile& tile,
                                         M64Lds<real, NW>& S) {
  constexpr bool PX = (PMASK & 1) != 0, PY = (PMASK & 2) != 0, PZ = (PMASK & 4) != 0;
  constexpr bool PML = PMASK != 0;
  constexpr uint32_t B = sizeof(real);
  auto& sEx = S.ex;
  auto& sEz = S.ez;
  auto& sHx = S.hx;
  auto& sHz = S.hz;
  auto& sX = S.x;
  const MarchArgs<real>& a = BATCH ? arg.sims[tile.pad0] : arg;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nyz = (size_t)ny * nz;
  const size_t fplane = (size_t)ny * a.zrow, fcell = (size_t)nx * fplane;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int jn = tile.jn, kn = tile.kn, k0 = tile.k0;
  const int j = tile.j0 - 1 + w, k = k0 + l;
  const bool row_ok = j >= 0 && j < ny && w <= jn + 1;
  const bool valid = row_ok && l < kn;
  const bool hrole = valid && w <= jn;
  const bool own = hrole && w >= 1;
  const int jc = min(max(j, 0), ny - 1), kc = min(k, nz - 1);
  const int i0 = tile.i0, i1 = tile.i1;
  const int ib = i0 > 0 ? i0 - 1 : 0;
  // rows whose stores cover whole lines: lanes past kn (only on a row's last
  // tile) write slots past nz
  const int kn_store = min(64, (kn + 31) & ~31);
  const bool st_lane = j >= 0 && j < ny && w >= 1 && w <= jn && l < kn_store;
  const bool lo_ok = k0 > 0;            // an apron node below the tile exists
  const bool up_ok = k0 + kn < nz;      // and above it

  // Buffer resources made once per tile, one for the E and one for the H
  // components of each level (component and plane offsets go in the scalar
  // operand), one for ce and one per psi side. The host runs this kernel
  // only while three field components, and each psi side, span less than
  // 2^31 bytes, so every scalar offset stays below 2^31.
  const uint32_t fcB = (uint32_t)(fcell * B), fpB = (uint32_t)(fplane * B);
  const __amdgpu_buffer_rsrc_t rEi = buf_rsrc(a.Lin), rHi = buf_rsrc(a.Lin + 3 * fcell);
  const __amdgpu_buffer_rsrc_t rEo = buf_rsrc(a.Lout), rHo = buf_rsrc(a.Lout + 3 * fcell);
  const __amdgpu_buffer_rsrc_t rCe = buf_rsrc(a.ce), rPE = buf_rsrc(a.psiE),
                               rPHi = buf_rsrc(a.psiHin), rPHo = buf_rsrc(a.psiHout);
  const uint32_t c1 = fcB, c2 = 2 * fcB;
  const size_t sy_e = (size_t)nx * a.pe[1] * nz, sz_e = (size_t)nx * ny * a.pe[2],
               sx_e = (size_t)a.pe[0] * nyz;
  const size_t sy_h = (size_t)nx * a.ph[1] * nz, sz_h = (size_t)nx * ny * a.ph[2],
               sx_h = (size_t)a.ph[0] * nyz;
  // byte offsets of the psi slots in their side's allocation:
  // E side ExY ExZ EyZ EyX EzX EzY, H side HxY HxZ HyZ HyX HzX HzY
  const uint32_t eoB[6] = {0, (uint32_t)(sy_e * B), (uint32_t)((sy_e + sz_e) * B),
                           (uint32_t)((sy_e + 2 * sz_e) * B), (uint32_t)((sy_e + 2 * sz_e + sx_e) * B),
                           (uint32_t)((sy_e + 2 * sz_e + 2 * sx_e) * B)};
  const size_t hoff[6] = {0, sy_h, sy_h + sz_h, sy_h + 2 * sz_h, sy_h + 2 * sz_h + sx_h,
                          sy_h + 2 * sz_h + 2 * sx_h};
  const uint32_t hoB[6] = {0, (uint32_t)(hoff[1] * B), (uint32_t)(hoff[2] * B), (uint32_t)(hoff[3] * B),
                           (uint32_t)(hoff[4] * B), (uint32_t)(hoff[5] * B)};
  const real* coef_y = a.coef + 4 * (size_t)nx;
  const real* coef_z = a.coef + 4 * (size_t)(nx + ny);

  for (int t = tid; PX && t < 4 * (i1 - ib); t += NW * 64) {
    const int q = t / (i1 - ib), ii = ib + t % (i1 - ib);
    sX[q][ii - ib] = a.coef[(size_t)q * nx + ii];
  }

  const Slabs<real> ye{a.slab_e[1][0], a.slab_e[1][1]}, yh{a.slab_h[1][0], a.slab_h[1][1]};
  const Slabs<real> ze{a.slab_e[2][0], a.slab_e[2][1]}, zh{a.slab_h[2][0], a.slab_h[2][1]};
  const Slabs<real> xe{a.slab_e[0][0], a.slab_e[0][1]}, xh{a.slab_h[0][0], a.slab_h[0][1]};
  const bool pyh = PY && hrole && yh.in(jc), pzh = PZ && hrole && zh.in(kc);
  const bool pye = PY && own && ye.in(jc), pze = PZ && own && ze.in(kc);
  const real bye = pye ? coef_y[jc] : real(0), cye = pye ? coef_y[ny + jc] : real(0);
  const real byh = pyh ? coef_y[2 * ny + jc] : real(0), cyh = pyh ? coef_y[3 * ny + jc] : real(0);
  const real bze = pze ? coef_z[kc] : real(0), cze = pze ? coef_z[nz + kc] : real(0);
  const real bzh = pzh ? coef_z[2 * nz + kc] : real(0), czh = pzh ? coef_z[3 * nz + kc] : real(0);
  // the apron node k0-1 of this row (uniform per wave)
  const int ka = k0 - 1;
  const bool row_h = j >= 0 && j < ny && w <= jn;       // rows that update H
  const bool row_e = row_h && w >= 1;                    // rows that update E
  const bool pyh_a = PY && row_e && lo_ok && yh.in(jc);
  const bool pzh_a = PZ && row_e && lo_ok && zh.in(ka);
  const real byh_a = pyh_a ? coef_y[2 * ny + jc] : real(0), cyh_a = pyh_a ? coef_y[3 * ny + jc] : real(0);
  const real bzh_a = pzh_a ? coef_z[2 * nz + ka] : real(0), czh_a = pzh_a ? coef_z[3 * nz + ka] : real(0);

  const uint32_t col = (uint32_t)(jc * nz + kc) * B;          // psi, dense rows of nz
  // fields / ce, rows of zrow: the lane's slot
  const uint32_t colf = (uint32_t)(jc * a.zrow + k) * B;
  const uint32_t cyh_o = (uint32_t)(yh.m(jc) * nz + kc) * B;
  const uint32_t czh_o = (uint32_t)(jc * a.ph[2] + zh.m(kc)) * B;
  const uint32_t cye_o = (uint32_t)(ye.m(jc) * nz + kc) * B;
  const uint32_t cze_o = (uint32_t)(jc * a.pe[2] + ze.m(kc)) * B;
  const size_t ys_h = (size_t)a.ph[1] * nz, zs_h = (size_t)ny * a.ph[2];
  const size_t ys_e = (size_t)a.pe[1] * nz, zs_e = (size_t)ny * a.pe[2];
  // per-plane byte strides of the y- and z-slab psi slots, of an x-slab psi plane
  const uint32_t ysB_h = (uint32_t)(ys_h * B), zsB_h = (uint32_t)(zs_h * B);
  const uint32_t ysB_e = (uint32_t)(ys_e * B), zsB_e = (uint32_t)(zs_e * B);
  const uint32_t xsB = (uint32_t)(nyz * B);

  // Lane offsets of every access, BUF_NONE where the lane does not take part
  // (the access is issued on every path; see BUF_NONE).
  const uint32_t f_l = valid ? colf : BUF_NONE;       // E loads
  const uint32_t f_h = hrole ? colf : BUF_NONE;       // H loads
  const uint32_t f_ce = own ? colf : BUF_NONE;
  // field stores: rows of a row's last tile are rounded up to 32 lanes (slots
  // past nz), so a row store covers whole lines
  const uint32_t f_st = st_lane ? colf : BUF_NONE;
  const uint32_t o_pyh = pyh ? cyh_o : BUF_NONE, o_pzh = pzh ? czh_o : BUF_NONE;
  const uint32_t o_pye = pye ? cye_o : BUF_NONE, o_pze = pze ? cze_o : BUF_NONE;
  const uint32_t o_h = hrole ? col : BUF_NONE, o_e = own ? col : BUF_NONE;
  const bool yh_in = j < ny - 1, zh_in = k < nz - 1;
  const bool ye_in = j >= 1, ze_in = k >= 1;

  // Apron gather: lane q < A_COUNT of a row that updates H fetches slot q
  // (E slots from plane ie, H and psi slots from plane ih). Slot pointer at
  // plane 0 and per-plane stride are fixed per lane; the x-slab psi slot
  // steps by slab plane. One unconditional global load per plane; a lane
  // whose slot does not exist reads the first field element instead. Its
  // value is never used: a slot is read only where it exists (E above the
  // tile only at k < nz-1, the slots below only for k0 > 0 on rows that
  // update E, row j+1 only for j < ny-1, plane i+1 only for i < nx-1, psi
  // only inside its slab). Masking it after the load would make the load's
  // consumer wait for every outstanding load.
  const int q = l;
  const bool eslot = q < A_HX_LO, xq = PX && q == A_PHYX;
  const real* a_base;
  size_t a_stride;
  bool a_ok;
  {
    const int jup = min(j + 1, ny - 1);
    const int kup = min(k0 + kn, nz - 1), klo = max(ka, 0);
    const bool lo_slot = lo_ok && row_e;
    a_ok = row_h && q < (PML ? (int)A_COUNT : (int)A_PHXY);
    a_ok = a_ok && (q <= A_EY_UP ? up_ok : lo_slot);
    a_ok = a_ok && (q != A_EZ_LO_JP || j + 1 < ny);
    // fields: component Ex Ey Ex Ey Ez Ez Hx Hy
    const int comp = q <= A_EY_UP ? q : q <= A_EY_LO ? q - 2 : q <= A_EZ_LO_JP ? 2 : q - 3;
    const int row = q == A_EZ_LO_JP ? jup : jc;
    const int kk = q <= A_EY_UP ? kup : klo;
    a_base = a.Lin + (size_t)min(comp, 5) * fcell + (size_t)row * a.zrow + kk;
    a_stride = fplane;
    if (PML && q >= A_PHXY && q < A_COUNT) {
      // psi slots: HxY (y slab), HxZ / HyZ (z slab), HyX (x slab)
      const bool yq = q == A_PHXY;
      a_ok = a_ok && (yq ? pyh_a : q == A_PHYX ? PX : pzh_a);
      a_base = a.psiHin + hoff[yq ? 0 : xq ? 3 : q == A_PHXZ ? 1 : 2] +
               (yq ? (size_t)yh.m(jc) * nz + klo : xq ? (size_t)jc * nz + klo
                                                    : (size_t)jc * a.ph[2] + zh.m(klo));
      a_stride = yq ? ys_h : xq ? nyz : zs_h;
    }
  }
  auto apron = [&](int ie, int ih, bool hslots) __attribute__((always_inline)) -> real {
    const bool hx_ok = !xq || xh.in(ih);
    const bool ok = a_ok && (eslot ? ie < nx : hslots && ih < nx && hx_ok);
    const int idx = eslot ? ie : xq ? xh.m(ih) : ih;
    const real* p = ok ? a_base + (size_t)idx * a_stride : a.Lin;
    return *gptr(p);
  };

  struct Plane {
    real e[3];    // E^n(i+1)
    real h[3];    // H^{n-1/2}(i)
    real ce;
    real ph[6], pe[6];
    real apr;     // apron gather
  };
  auto fetch = [&](Plane& P, int i) __attribute__((always_inline)) {
    const uint32_t so = (uint32_t)i * fpB;
    const uint32_t son = (uint32_t)min(i + 1, nx - 1) * fpB;
    const uint32_t oe = i + 1 < nx ? f_l : BUF_NONE;
    P.e[0] = blds<real>(rEi, oe, son);
    P.e[1] = blds<real>(rEi, oe, son + c1);
    P.e[2] = blds<real>(rEi, oe, son + c2);
    P.h[0] = blds<real>(rHi, f_h, so);
    P.h[1] = blds<real>(rHi, f_h, so + c1);
    P.h[2] = blds<real>(rHi, f_h, so + c2);
    P.ce = blds<real>(rCe, f_ce, so);
    P.apr = apron(i + 1, i, true);
    for (int c = 0; c < 6; ++c) P.ph[c] = P.pe[c] = real(0);
    if constexpr (PY) {
      const uint32_t yh = (uint32_t)i * ysB_h, ye = (uint32_t)i * ysB_e;
      P.ph[0] = blds<real>(rPHi, o_pyh, hoB[0] + yh);
      P.ph[5] = blds<real>(rPHi, o_pyh, hoB[5] + yh);
      P.pe[0] = blds<real>(rPE, o_pye, eoB[0] + ye);
      P.pe[5] = blds<real>(rPE, o_pye, eoB[5] + ye);
    }
    if constexpr (PZ) {
      const uint32_t zh = (uint32_t)i * zsB_h, ze = (uint32_t)i * zsB_e;
      P.ph[1] = blds<real>(rPHi, o_pzh, hoB[1] + zh);
      P.ph[2] = blds<real>(rPHi, o_pzh, hoB[2] + zh);
      P.pe[1] = blds<real>(rPE, o_pze, eoB[1] + ze);
      P.pe[2] = blds<real>(rPE, o_pze, eoB[2] + ze);
    }
    if constexpr (PX) {
      const bool xhi = xh.in(i), xei = xe.in(i);
      const uint32_t xoh = (uint32_t)(xhi ? xh.m(i) : 0) * xsB, xoe = (uint32_t)(xei ? xe.m(i) : 0) * xsB;
      P.ph[3] = blds<real>(rPHi, xhi ? o_h : BUF_NONE, hoB[3] + xoh);
      P.ph[4] = blds<real>(rPHi, xhi ? o_h : BUF_NONE, hoB[4] + xoh);
      P.pe[3] = blds<real>(rPE, xei ? o_e : BUF_NONE, eoB[3] + xoe);
      P.pe[4] = blds<real>(rPE, xei ? o_e : BUF_NONE, eoB[4] + xoe);
    }
  };

  // E^n(ib) and its apron (E slots at plane ib)
  real ecur[3];
  {
    const uint32_t so = (uint32_t)ib * fpB;
    ecur[0] = blds<real>(rEi, f_l, so);
    ecur[1] = blds<real>(rEi, f_l, so + c1);
    ecur[2] = blds<real>(rEi, f_l, so + c2);
  }
  real acur = apron(ib, ib, false);
  real hy_prev = 0, hz_prev = 0;
  const real K_h = -a.ch;

  auto plane = [&](int i, const Plane& P, Plane& Next) __attribute__((always_inline)) {
    if (i + DIST < i1) fetch(Next, i + DIST);
    const bool live = i >= i0;
    const int bi = i & 1;
    const uint32_t so = (uint32_t)i * fpB;
    const real ex_jp = w < NW - 1 ? sEx[bi][min(tid + 64, NW * 64 - 1)] : real(0);
    const real ez_jp = w < NW - 1 ? sEz[bi][min(tid + 64, NW * 64 - 1)] : real(0);
    // E(k+1): the next lane, the apron above the tile for the last node
    const real ex_up = lane_value(acur, A_EX_UP), ey_up = lane_value(acur, A_EY_UP);
    real ex_kp = lane_next(ecur[0]);
    real ey_kp = lane_next(ecur[1]);
    if (l == kn - 1) {
      ex_kp = ex_up;
      ey_kp = ey_up;
    }
    // ---- H^{n+1/2}(i) ----
    const bool xh_in = i < nx - 1;
    const bool pxh = PX && xh.in(i);
    const real bxh = PX ? sX[2][i - ib] : real(0), cxh = PX ? sX[3][i - ib] : real(0);
    real hx = P.h[0], hy = P.h[1], hz = P.h[2];
    real nps[6] = {0, 0, 0, 0, 0, 0};
    if (hrole) {
      if (yh_in && zh_in) {
        const real dA = dd<real, EXACT>(ez_jp, ecur[2], a.d[1], a.inv_d[1]);
        const real dB = dd<real, EXACT>(ey_kp, ecur[1], a.d[2], a.inv_d[2]);
        hx = curl_update(hx, K_h, true, dA, pyh, P.ph[0], byh, cyh, &nps[0], true, dB, pzh,
                         P.ph[1], bzh, czh, &nps[1]);
      }
      if (zh_in && xh_in) {
        const real dA = dd<real, EXACT>(ex_kp, ecur[0], a.d[2], a.inv_d[2]);
        const real dB = dd<real, EXACT>(P.e[2], ecur[2], a.d[0], a.inv_d[0]);
        hy = curl_update(hy, K_h, true, dA, pzh, P.ph[2], bzh, czh, &nps[2], true, dB, pxh,
                         P.ph[3], bxh, cxh, &nps[3]);
      }
      if (xh_in && yh_in) {
        const real dA = dd<real, EXACT>(P.e[1], ecur[1], a.d[0], a.inv_d[0]);
        const real dB = dd<real, EXACT>(ex_jp, ecur[0], a.d[1], a.inv_d[1]);
        hz = curl_update(hz, K_h, true, dA, pxh, P.ph[4], bxh, cxh, &nps[4], true, dB, pyh,
                         P.ph[5], byh, cyh, &nps[5]);
      }
    }
    if constexpr (PML) {
      // H-side psi of the owned nodes, where the curl term exists
      const bool oyz = own && live && yh_in && zh_in;
      const bool ozx = own && live && zh_in && xh_in;
      const bool oxy = own && live && xh_in && yh_in;
      if constexpr (PY) {
        const uint32_t yh = (uint32_t)i * ysB_h;
        bsts(rPHo, oyz ? o_pyh : BUF_NONE, hoB[0] + yh, nps[0]);
        bsts(rPHo, oxy ? o_pyh : BUF_NONE, hoB[5] + yh, nps[5]);
      }
      if constexpr (PZ) {
        const uint32_t zh = (uint32_t)i * zsB_h;
        bsts(rPHo, oyz ? o_pzh : BUF_NONE, hoB[1] + zh, nps[1]);
        bsts(rPHo, ozx ? o_pzh : BUF_NONE, hoB[2] + zh, nps[2]);
      }
      if constexpr (PX) {
        const uint32_t xo = (uint32_t)(pxh ? xh.m(i) : 0) * xsB;
        bsts(rPHo, ozx && pxh ? o_h : BUF_NONE, hoB[3] + xo, nps[3]);
        bsts(rPHo, oxy && pxh ? o_h : BUF_NONE, hoB[4] + xo, nps[4]);
      }
    }
    // H^{n+1/2}(i) at k0-1, recomputed as its owner does (wave-uniform)
    real hx_a = 0, hy_a = 0;
    if (lo_ok && row_e) {
      const real ex_lo = lane_value(acur, A_EX_LO), ey_lo = lane_value(acur, A_EY_LO);
      const real ez_lo = lane_value(acur, A_EZ_LO), ez_lo_jp = lane_value(acur, A_EZ_LO_JP);
      const real ez_lo_ip = lane_value(P.apr, A_EZ_LO);
      const real ex_k0 = lane_value(ecur[0], 0), ey_k0 = lane_value(ecur[1], 0);
      hx_a = lane_value(P.apr, A_HX_LO);
      hy_a = lane_value(P.apr, A_HY_LO);
      real dummy0 = 0, dummy1 = 0;
      if (yh_in) {   // z: ka < nz - 1 always
        const real dA = dd<real, EXACT>(ez_lo_jp, ez_lo, a.d[1], a.inv_d[1]);
        const real dB = dd<real, EXACT>(ey_k0, ey_lo, a.d[2], a.inv_d[2]);
        hx_a = curl_update(hx_a, K_h, true, dA, pyh_a, lane_value(P.apr, A_PHXY), byh_a, cyh_a,
                           &dummy0, true, dB, pzh_a, lane_value(P.apr, A_PHXZ), bzh_a, czh_a,
                           &dummy1);
      }
      if (xh_in) {
        const real dA = dd<real, EXACT>(ex_k0, ex_lo, a.d[2], a.inv_d[2]);
        const real dB = dd<real, EXACT>(ez_lo_ip, ez_lo, a.d[0], a.inv_d[0]);
        hy_a = curl_update(hy_a, K_h, true, dA, pzh_a, lane_value(P.apr, A_PHYZ), bzh_a, czh_a,
                           &dummy0, true, dB, pxh, lane_value(P.apr, A_PHYX), bxh, cxh,
                           &dummy1);
      }
    }
    const uint32_t fs = live ? f_st : BUF_NONE;
    bsts<BUF_NT>(rHo, fs, so, hx);
    bsts<BUF_NT>(rHo, fs, so + c1, hy);
    bsts<BUF_NT>(rHo, fs, so + c2, hz);
    sHx[bi][tid] = hx;
    sHz[bi][tid] = hz;
    sEx[bi ^ 1][tid] = P.e[0];
    sEz[bi ^ 1][tid] = P.e[2];
    lds_barrier();
    const real hx_jm = w > 0 ? sHx[bi][tid - 64] : real(0);
    const real hz_jm = w > 0 ? sHz[bi][tid - 64] : real(0);
    real hx_km = lane_prev(hx);
    real hy_km = lane_prev(hy);
    if (l == 0) {
      hx_km = hx_a;
      hy_km = hy_a;
    }
    // ---- E^{n+1}(i) ----
    const bool xe_in = i >= 1;
    const bool pxe = PX && xe.in(i);
    real ex = ecur[0], ey = ecur[1], ez = ecur[2];
    real npe[6] = {0, 0, 0, 0, 0, 0};
    if (st_lane) {
      const real bxe = PX ? sX[0][i - ib] : real(0), cxe = PX ? sX[1][i - ib] : real(0);
      const real K = P.ce;
      if (ye_in && ze_in) {
        const real dA = dd<real, EXACT>(hz, hz_jm, a.d[1], a.inv_d[1]);
        const real dB = dd<real, EXACT>(hy, hy_km, a.d[2], a.inv_d[2]);
        ex = curl_update(ex, K, true, dA, pye, P.pe[0], bye, cye, &npe[0], true, dB, pze,
                         P.pe[1], bze, cze, &npe[1]);
      }
      if (ze_in && xe_in) {
        const real dA = dd<real, EXACT>(hx, hx_km, a.d[2], a.inv_d[2]);
        const real dB = dd<real, EXACT>(hz, hz_prev, a.d[0], a.inv_d[0]);
        ey = curl_update(ey, K, true, dA, pze, P.pe[2], bze, cze, &npe[2], true, dB, pxe,
                         P.pe[3], bxe, cxe, &npe[3]);
      }
      if (xe_in && ye_in) {
        const real dA = dd<real, EXACT>(hy, hy_prev, a.d[0], a.inv_d[0]);
        const real dB = dd<real, EXACT>(hx, hx_jm, a.d[1], a.inv_d[1]);
        ez = curl_update(ez, K, true, dA, pxe, P.pe[4], bxe, cxe, &npe[4], true, dB, pye,
                         P.pe[5], bye, cye, &npe[5]);
      }
    }
    bsts<BUF_NT>(rEo, fs, so, ex);
    bsts<BUF_NT>(rEo, fs, so + c1, ey);
    bsts<BUF_NT>(rEo, fs, so + c2, ez);
    if constexpr (PML) {
      const bool oyz = own && live && ye_in && ze_in;
      const bool ozx = own && live && ze_in && xe_in;
      const bool oxy = own && live && xe_in && ye_in;
      if constexpr (PY) {
        const uint32_t ye = (uint32_t)i * ysB_e;
        bsts(rPE, oyz ? o_pye : BUF_NONE, eoB[0] + ye, npe[0]);
        bsts(rPE, oxy ? o_pye : BUF_NONE, eoB[5] + ye, npe[5]);
      }
      if constexpr (PZ) {
        const uint32_t ze = (uint32_t)i * zsB_e;
        bsts(rPE, oyz ? o_pze : BUF_NONE, eoB[1] + ze, npe[1]);
        bsts(rPE, ozx ? o_pze : BUF_NONE, eoB[2] + ze, npe[2]);
      }
      if constexpr (PX) {
        const uint32_t xo = (uint32_t)(pxe ? xe.m(i) : 0) * xsB;
        bsts(rPE, ozx && pxe ? o_e : BUF_NONE, eoB[3] + xo, npe[3]);
        bsts(rPE, oxy && pxe ? o_e : BUF_NONE, eoB[4] + xo, npe[4]);
      }
    }
    hy_prev = hy;
    hz_prev = hz;
  };
  sEx[ib & 1][tid] = ecur[0];
  sEz[ib & 1][tid] = ecur[2];
  lds_barrier();
  if constexpr (DIST == 1) {
    // two plane buffers alternate
    Plane A, Bq;
    fetch(A, ib);
    for (int i = ib; i < i1; i += 2) {
      if (i > ib) {
        ecur[0] = Bq.e[0];
        ecur[1] = Bq.e[1];
        ecur[2] = Bq.e[2];
        acur = Bq.apr;
      }
      plane(i, A, Bq);
      if (i + 1 >= i1) break;
      ecur[0] = A.e[0];
      ecur[1] = A.e[1];
      ecur[2] = A.e[2];
      acur = A.apr;
      plane(i + 1, Bq, A);
    }
    return;
  }
  // Three plane buffers rotate: plane i is computed from one while planes
  // i + 1 and i + 2 are in flight in the others; E^n(i) and its apron are
  // copied out of the buffer of plane i - 1 before that buffer is refilled.
  Plane A, Bq, Cq;
  fetch(A, ib);
  if (ib + 1 < i1) fetch(Bq, ib + 1);
  for (int i = ib; i < i1; i += 3) {
    if (i > ib) {
      ecur[0] = Cq.e[0];
      ecur[1] = Cq.e[1];
      ecur[2] = Cq.e[2];
      acur = Cq.apr;
    }
    plane(i, A, Cq);
    if (i + 1 >= i1) break;
    ecur[0] = A.e[0];
    ecur[1] = A.e[1];
    ecur[2] = A.e[2];
    acur = A.apr;
    plane(i + 1, Bq, A);
    if (i + 2 >= i1) break;
    ecur[0] = Bq.e[0];
    ecur[1] = Bq.e[1];
    ecur[2] = Bq.e[2];
    acur = Bq.apr;
    plane(i + 2, Cq, Bq);
  }
}

// PMASK: the CPML axes the tile's nodes may lie in (bit 0 x, 1 y, 2 z); the
// psi terms of the other axes are left out at compile time. DIST: planes of
// load prefetch (2: three rotating plane buffers; 1: two).
// float32 variants other than the all-axis one are held to 80 VGPRs (six
// waves per SIMD: two 12-wave workgroups per CU).
template <typename real, bool EXACT, int NW, int PMASK, int DIST, bool BATCH = false>
__global__ __launch_bounds__(NW * 64, sizeof(real) == 4 && PMASK != 7 ? 6 : 1) void k_yee_m64(
    MarchArgs<real> arg) {
  __shared__ M64Lds<real, NW> S;
  if ((int)blockIdx.x >= arg.ntiles) return;
  const MarchTile tile = reinterpret_cast<const MarchTile*>(arg.tiles)[blockIdx.x];
  if (tile.i1 <= tile.i0) return;
  m64_tile<real, EXACT, NW, PMASK, DIST, BATCH>(arg, tile, S);
}

// Tiles of several CPML axis sets in one launch, one workgroup per tile, the
// set of each tile in its pad1 field (0: psi-free, 4: z slabs only). The
// list is in spatial order (z tiles of a row block next to each other), so
// a CPML tile and its psi-free z neighbours run at the same time on one XCD
// and share the z apron lines in its L2.
template <typename real, bool EXACT, int NW, int SETS>
__global__ __launch_bounds__(NW * 64, sizeof(real) == 4 ? 6 : 1) void k_yee_m64_mix(
    MarchArgs<real> arg) {
  __shared__ M64Lds<real, NW> S;
  if ((int)blockIdx.x >= arg.ntiles) return;
  const MarchTile tile = reinterpret_cast<const MarchTile*>(arg.tiles)[blockIdx.x];
  if (tile.i1 <= tile.i0) return;
  if (tile.pad1 == 0) m64_tile<real, EXACT, NW, 0, 2, false>(arg, tile, S);
  else if (tile.pad1 == 4) m64_tile<real, EXACT, NW, 4, 2, false>(arg, tile, S);
  else if ((SETS & 2) && tile.pad1 == 2) m64_tile<real, EXACT, NW, 2, 2, false>(arg, tile, S);
  else if (SETS & 1) m64_tile<real, EXACT, NW, 1, 2, false>(arg, tile, S);
}

// Variants: psi-free and single-axis CPML tiles prefetch two planes (three
// buffers; they fit two 12-wave workgroups per CU), the all-axis CPML
// variant one (two buffers; one workgroup per CU within 168 VGPRs).
template <typename real, bool EXACT, int NW, bool BATCH>
void launch_m64_mask(MarchArgs<real>& a, int mask, hipStream_t s) {
  const dim3 g(a.ntiles), b(NW * 64);
  switch (mask) {
    case 0: k_yee_m64<real, EXACT, NW, 0, 2, BATCH><<<g, b, 0, s>>>(a); return;
    case 7: k_yee_m64<real, EXACT, NW, 7, 1, BATCH><<<g, b, 0, s>>>(a); return;
    default: break;
  }
  if constexpr (!BATCH) {
    switch (mask) {
      case 1: k_yee_m64<real, EXACT, NW, 1, 2><<<g, b, 0, s>>>(a); return;
      case 2: k_yee_m64<real, EXACT, NW, 2, 2><<<g, b, 0, s>>>(a); return;
      case 4: k_yee_m64<real, EXACT, NW, 4, 2><<<g, b, 0, s>>>(a); return;
      default: break;
    }
  }
  throw std::runtime_error("march64: no kernel for this CPML axis set");
}

template <typename real, int NW>
void launch_m64_nw(MarchArgs<real>& a, bool exact, int mask, hipStream_t s) {
  if (a.sims) {
    if (exact) launch_m64_mask<real, true, NW, true>(a, mask, s);
    else launch_m64_mask<real, false, NW, true>(a, mask, s);
  } else {
    if (exact) launch_m64_mask<real, true, NW, false>(a, mask, s);
    else launch_m64_mask<real, false, NW, false>(a, mask, s);
  }
}

template <typename real>
void launch_march64(MarchArgs<real>& a, bool exact, int mask, int nw, hipStream_t s) {
  if (a.ntiles == 0) return;
  if (nw == 16) launch_m64_nw<real, 16>(a, exact, mask, s);
  else if (nw == 12) launch_m64_nw<real, 12>(a, exact, mask, s);
  else launch_m64_nw<real, MARCH_NW>(a, exact, mask, s);
}

template void launch_march64<float>(MarchArgs<float>&, bool, int, int, hipStream_t);
template void launch_march64<double>(MarchArgs<double>&, bool, int, int, hipStream_t);

template <typename real>
void launch_march64_mix(MarchArgs<real>& a, bool exact, int nw, int sets, hipStream_t s) {
  if (a.ntiles == 0) return;
  if (nw != 12) throw std::runtime_error("march64: the mixed launch runs 12-wave workgroups");
  const dim3 g(a.ntiles), b(12 * 64);
  // sets besides psi-free and z-only: bit 1 y-only, bit 0 x-only
  switch (sets) {
    case 0: exact ? k_yee_m64_mix<real, true, 12, 0><<<g, b, 0, s>>>(a) : k_yee_m64_mix<real, false, 12, 0><<<g, b, 0, s>>>(a); return;
    case 2: exact ? k_yee_m64_mix<real, true, 12, 2><<<g, b, 0, s>>>(a) : k_yee_m64_mix<real, false, 12, 2><<<g, b, 0, s>>>(a); return;
    case 3: exact ? k_yee_m64_mix<real, true, 12, 3><<<g, b, 0, s>>>(a) : k_yee_m64_mix<real, false, 12, 3><<<g, b, 0, s>>>(a); return;
    default: break;
  }
  throw std::runtime_error("march64: no mixed kernel for these CPML axis sets");
}

template void launch_march64_mix<float>(MarchArgs<float>&, bool, int, int, hipStream_t);
template void launch_march64_mix<double>(MarchArgs<double>&, bool, int, int, hipStream_t);

}  // namespace pfhip
