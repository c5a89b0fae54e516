// Fused 3-D Yee + CPML step for gfx950: one launch advances H and E.
//
// Data layout in HBM: six C-order (nx, ny, nz) field arrays per time level,
// z fastest, ping-ponged (E^n, H^{n-1/2} in; E^{n+1}, H^{n+1/2} out) so that
// workgroups may recompute their neighbours' H apron without racing the
// owner's writes; the H-side CPML psi is ping-ponged for the same reason.
//
// Work decomposition: a workgroup owns a TY x 64 (y, z) tile -- one
// wavefront per y row, lanes along z -- and marches along x over a chunk of
// planes. Per plane i it stages E^n(i+1) (tile + one-node apron) in LDS, forms
// H^{n+1/2}(i) on the tile plus the j0-1 / k0-1 aprons (LDS), then E^{n+1}(i)
// on the tile from H(i), H(i-1) (registers) and E^n(i). Algorithmic traffic
// per node and step: read 6 fields + ce, write 6 fields (13 words).
//
// Per-node arithmetic is the reference order (simulation.py:1252-1309) when
// EXACT (difference / cell size, then bulk-then-slab PML terms), so float64
// runs stay bit-identical to the generic kernels; otherwise differences are
// scaled by 1/cell size.
#include "pfhip_internal.h"
#include "yee_common.h"

namespace pfhip {

namespace {

constexpr int TZ = 64;

template <typename real, bool EXACT>
__device__ __forceinline__ real ddiff(real hi, real lo, real d, real inv_d) {
  if constexpr (EXACT) return (hi - lo) / d;
  else return (hi - lo) * inv_d;
}

// psi index helpers for the three compact layouts.
__device__ __forceinline__ int64_t pidx_x(int m, int j, int k, int ny, int nz) {
  return ((int64_t)m * ny + j) * nz + k;
}
__device__ __forceinline__ int64_t pidx_y(int i, int m, int k, int py, int nz) {
  return ((int64_t)i * py + m) * nz + k;
}
__device__ __forceinline__ int64_t pidx_z(int i, int j, int m, int ny, int pz) {
  return ((int64_t)i * ny + j) * pz + m;
}


}  // namespace

template <typename real, bool EXACT, int TY>
__global__ __launch_bounds__(TY * 64) void k_yee_fused(FusedArgs<real> a) {
  constexpr int NT = TY * 64;
  constexpr int EW = TZ + 2, EH = TY + 2;   // E tile: rows j0-1..j0+TY, cols k0-1..k0+TZ
  constexpr int HW = TZ + 1, HH = TY + 1;   // H tile: rows j0-1..j0+TY-1, cols k0-1..k0+TZ-1
  __shared__ real sE[2][3][EH * EW];
  __shared__ real sH[3][HH * HW];

  // XCD-aware order: consecutive blocks of one XCD take neighbouring tiles.
  const int nblk = a.tiles_z * a.tiles_y * a.chunks_x;
  const int per_xcd = a.grid_blocks / 8;
  const int bid = blockIdx.x;
  const int logical = (bid % 8) * per_xcd + bid / 8;
  if (logical >= nblk) return;
  const int tzi = logical % a.tiles_z;
  const int tyi = (logical / a.tiles_z) % a.tiles_y;
  const int cxi = logical / (a.tiles_z * a.tiles_y);

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t nyz = a.nyz;
  const int j0 = tyi * TY, k0 = tzi * TZ;
  const int i0 = cxi * a.xchunk;
  const int i1 = min(nx, i0 + a.xchunk);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wy = tid >> 6;
  const int j = j0 + wy, k = k0 + lane;
  const bool own = j < ny && k < nz;
  const int own_e = (wy + 1) * EW + (lane + 1);
  const int own_h = (wy + 1) * HW + (lane + 1);

  const real* __restrict__ Ein0 = a.Ein[0];
  const real* __restrict__ Ein1 = a.Ein[1];
  const real* __restrict__ Ein2 = a.Ein[2];

  auto load_e = [&](int buf, int i) {
    for (int idx = tid; idx < EH * EW; idx += NT) {
      const int r = idx / EW, c = idx - (idx / EW) * EW;
      const int jj = j0 - 1 + r, kk = k0 - 1 + c;
      real v0 = 0, v1 = 0, v2 = 0;
      if (i < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz) {
        const int64_t g = (int64_t)i * nyz + (int64_t)jj * nz + kk;
        v0 = Ein0[g];
        v1 = Ein1[g];
        v2 = Ein2[g];
      }
      sE[buf][0][idx] = v0;
      sE[buf][1][idx] = v1;
      sE[buf][2][idx] = v2;
    }
  };

  // H^{n+1/2} on plane i over the tile + apron; owners store H and psi.
  auto h_plane = [&](int i, int cur, int nxt, bool store) {
    const real K = -a.ch;
    const int mx = a.map_h[0][i];
    const bool xin = i < nx - 1;
    for (int idx = tid; idx < HH * HW; idx += NT) {
      const int r = idx / HW, c = idx - (idx / HW) * HW;
      const int jj = j0 - 1 + r, kk = k0 - 1 + c;
      real hx = 0, hy = 0, hz = 0;
      if (jj >= 0 && jj < ny && kk >= 0 && kk < nz) {
        const int64_t g = (int64_t)i * nyz + (int64_t)jj * nz + kk;
        const bool owner = store && r >= 1 && c >= 1;
        const int e = r * EW + c;                     // same node in the E tile
        const real* Ex = sE[cur][0];
        const real* Ey = sE[cur][1];
        const real* Ez = sE[cur][2];
        hx = a.Hin[0][g];
        hy = a.Hin[1][g];
        hz = a.Hin[2][g];
        const bool yin = jj < ny - 1, zin = kk < nz - 1;
        const int my = a.map_h[1][jj], mz = a.map_h[2][kk];
        // Hx -= ch (d_y Ez - d_z Ey)
        if (yin && zin) {
          const real dA = ddiff<real, EXACT>(Ez[e + EW], Ez[e], a.d[1], a.inv_d[1]);
          const real dB = ddiff<real, EXACT>(Ey[e + 1], Ey[e], a.d[2], a.inv_d[2]);
          real pa = 0, pb = 0, na = 0, nb = 0;
          if (my >= 0) pa = a.psiHin[PSI_HXY - 6][pidx_y(i, my, kk, a.ph[1], nz)];
          if (mz >= 0) pb = a.psiHin[PSI_HXZ - 6][pidx_z(i, jj, mz, ny, a.ph[2])];
          hx = curl_update(hx, K, true, dA, my >= 0, pa, my >= 0 ? a.b_h[1][jj] : real(0),
                           my >= 0 ? a.c_h[1][jj] : real(0), &na, true, dB, mz >= 0, pb,
                           mz >= 0 ? a.b_h[2][kk] : real(0), mz >= 0 ? a.c_h[2][kk] : real(0), &nb);
          if (owner) {
            if (my >= 0) a.psiHout[PSI_HXY - 6][pidx_y(i, my, kk, a.ph[1], nz)] = na;
            if (mz >= 0) a.psiHout[PSI_HXZ - 6][pidx_z(i, jj, mz, ny, a.ph[2])] = nb;
          }
        }
        // Hy -= ch (d_z Ex - d_x Ez)
        if (zin && xin) {
          const real dA = ddiff<real, EXACT>(Ex[e + 1], Ex[e], a.d[2], a.inv_d[2]);
          const real dB = ddiff<real, EXACT>(sE[nxt][2][e], Ez[e], a.d[0], a.inv_d[0]);
          real pa = 0, pb = 0, na = 0, nb = 0;
          if (mz >= 0) pa = a.psiHin[PSI_HYZ - 6][pidx_z(i, jj, mz, ny, a.ph[2])];
          if (mx >= 0) pb = a.psiHin[PSI_HYX - 6][pidx_x(mx, jj, kk, ny, nz)];
          hy = curl_update(hy, K, true, dA, mz >= 0, pa, mz >= 0 ? a.b_h[2][kk] : real(0),
                           mz >= 0 ? a.c_h[2][kk] : real(0), &na, true, dB, mx >= 0, pb,
                           mx >= 0 ? a.b_h[0][i] : real(0), mx >= 0 ? a.c_h[0][i] : real(0), &nb);
          if (owner) {
            if (mz >= 0) a.psiHout[PSI_HYZ - 6][pidx_z(i, jj, mz, ny, a.ph[2])] = na;
            if (mx >= 0) a.psiHout[PSI_HYX - 6][pidx_x(mx, jj, kk, ny, nz)] = nb;
          }
        }
        // Hz -= ch (d_x Ey - d_y Ex)
        if (xin && yin) {
          const real dA = ddiff<real, EXACT>(sE[nxt][1][e], Ey[e], a.d[0], a.inv_d[0]);
          const real dB = ddiff<real, EXACT>(Ex[e + EW], Ex[e], a.d[1], a.inv_d[1]);
          real pa = 0, pb = 0, na = 0, nb = 0;
          if (mx >= 0) pa = a.psiHin[PSI_HZX - 6][pidx_x(mx, jj, kk, ny, nz)];
          if (my >= 0) pb = a.psiHin[PSI_HZY - 6][pidx_y(i, my, kk, a.ph[1], nz)];
          hz = curl_update(hz, K, true, dA, mx >= 0, pa, mx >= 0 ? a.b_h[0][i] : real(0),
                           mx >= 0 ? a.c_h[0][i] : real(0), &na, true, dB, my >= 0, pb,
                           my >= 0 ? a.b_h[1][jj] : real(0), my >= 0 ? a.c_h[1][jj] : real(0), &nb);
          if (owner) {
            if (mx >= 0) a.psiHout[PSI_HZX - 6][pidx_x(mx, jj, kk, ny, nz)] = na;
            if (my >= 0) a.psiHout[PSI_HZY - 6][pidx_y(i, my, kk, a.ph[1], nz)] = nb;
          }
        }
        if (owner) {
          a.Hout[0][g] = hx;
          a.Hout[1][g] = hy;
          a.Hout[2][g] = hz;
        }
      }
      sH[0][idx] = hx;
      sH[1][idx] = hy;
      sH[2][idx] = hz;
    }
  };

  real hy_prev = 0, hz_prev = 0;   // H^{n+1/2}(i-1) at the own node
  int cur = 0;
  if (i0 > 0) {
    load_e(0, i0 - 1);
    load_e(1, i0);
    __syncthreads();
    h_plane(i0 - 1, 0, 1, false);
    __syncthreads();
    hy_prev = sH[1][own_h];
    hz_prev = sH[2][own_h];
    cur = 1;
  } else {
    load_e(0, i0);
  }

  for (int i = i0; i < i1; ++i) {
    const int nxt = cur ^ 1;
    load_e(nxt, i + 1);
    __syncthreads();
    h_plane(i, cur, nxt, true);
    __syncthreads();
    const real hy_i = sH[1][own_h];
    const real hz_i = sH[2][own_h];
    if (own) {
      const int64_t g = (int64_t)i * nyz + (int64_t)j * nz + k;
      const real K = a.ce[g];
      const real* Hx = sH[0];
      const real* Hy = sH[1];
      const real* Hz = sH[2];
      real ex = sE[cur][0][own_e], ey = sE[cur][1][own_e], ez = sE[cur][2][own_e];
      const bool xin = i >= 1, yin = j >= 1, zin = k >= 1;
      const int mx = a.map_e[0][i], my = a.map_e[1][j], mz = a.map_e[2][k];
      // Ex += ce (d_y Hz - d_z Hy)
      if (yin && zin) {
        const real dA = ddiff<real, EXACT>(Hz[own_h], Hz[own_h - HW], a.d[1], a.inv_d[1]);
        const real dB = ddiff<real, EXACT>(Hy[own_h], Hy[own_h - 1], a.d[2], a.inv_d[2]);
        real pa = 0, pb = 0, na = 0, nb = 0;
        if (my >= 0) pa = a.psiE[PSI_EXY][pidx_y(i, my, k, a.pe[1], nz)];
        if (mz >= 0) pb = a.psiE[PSI_EXZ][pidx_z(i, j, mz, ny, a.pe[2])];
        ex = curl_update(ex, K, true, dA, my >= 0, pa, my >= 0 ? a.b_e[1][j] : real(0),
                         my >= 0 ? a.c_e[1][j] : real(0), &na, true, dB, mz >= 0, pb,
                         mz >= 0 ? a.b_e[2][k] : real(0), mz >= 0 ? a.c_e[2][k] : real(0), &nb);
        if (my >= 0) a.psiE[PSI_EXY][pidx_y(i, my, k, a.pe[1], nz)] = na;
        if (mz >= 0) a.psiE[PSI_EXZ][pidx_z(i, j, mz, ny, a.pe[2])] = nb;
      }
      // Ey += ce (d_z Hx - d_x Hz)
      if (zin && xin) {
        const real dA = ddiff<real, EXACT>(Hx[own_h], Hx[own_h - 1], a.d[2], a.inv_d[2]);
        const real dB = ddiff<real, EXACT>(hz_i, hz_prev, a.d[0], a.inv_d[0]);
        real pa = 0, pb = 0, na = 0, nb = 0;
        if (mz >= 0) pa = a.psiE[PSI_EYZ][pidx_z(i, j, mz, ny, a.pe[2])];
        if (mx >= 0) pb = a.psiE[PSI_EYX][pidx_x(mx, j, k, ny, nz)];
        ey = curl_update(ey, K, true, dA, mz >= 0, pa, mz >= 0 ? a.b_e[2][k] : real(0),
                         mz >= 0 ? a.c_e[2][k] : real(0), &na, true, dB, mx >= 0, pb,
                         mx >= 0 ? a.b_e[0][i] : real(0), mx >= 0 ? a.c_e[0][i] : real(0), &nb);
        if (mz >= 0) a.psiE[PSI_EYZ][pidx_z(i, j, mz, ny, a.pe[2])] = na;
        if (mx >= 0) a.psiE[PSI_EYX][pidx_x(mx, j, k, ny, nz)] = nb;
      }
      // Ez += ce (d_x Hy - d_y Hx)
      if (xin && yin) {
        const real dA = ddiff<real, EXACT>(hy_i, hy_prev, a.d[0], a.inv_d[0]);
        const real dB = ddiff<real, EXACT>(Hx[own_h], Hx[own_h - HW], a.d[1], a.inv_d[1]);
        real pa = 0, pb = 0, na = 0, nb = 0;
        if (mx >= 0) pa = a.psiE[PSI_EZX][pidx_x(mx, j, k, ny, nz)];
        if (my >= 0) pb = a.psiE[PSI_EZY][pidx_y(i, my, k, a.pe[1], nz)];
        ez = curl_update(ez, K, true, dA, mx >= 0, pa, mx >= 0 ? a.b_e[0][i] : real(0),
                         mx >= 0 ? a.c_e[0][i] : real(0), &na, true, dB, my >= 0, pb,
                         my >= 0 ? a.b_e[1][j] : real(0), my >= 0 ? a.c_e[1][j] : real(0), &nb);
        if (mx >= 0) a.psiE[PSI_EZX][pidx_x(mx, j, k, ny, nz)] = na;
        if (my >= 0) a.psiE[PSI_EZY][pidx_y(i, my, k, a.pe[1], nz)] = nb;
      }
      a.Eout[0][g] = ex;
      a.Eout[1][g] = ey;
      a.Eout[2][g] = ez;
    }
    hy_prev = hy_i;
    hz_prev = hz_i;
    __syncthreads();
    cur = nxt;
  }
}

// Register-marching variant: no LDS and no barriers. Each thread owns one
// (j, k) column of a TY x 64 tile and marches x over a chunk of planes. H at
// its own node is computed once per plane; the H values its E update needs at
// (j-1) and (k-1) are recomputed from cached loads (they belong to the
// neighbouring threads), and H(i-1) is carried in registers. All global loads
// of a plane are independent, so the wavefront keeps many in flight.
template <typename real, bool EXACT>
struct HNode {
  // H^{n+1/2} components at node (i, jj, kk); which ones are needed is a
  // compile-time mask (1 = Hx, 2 = Hy, 4 = Hz). The owner writes H and psi.
  template <int MASK>
  __device__ __forceinline__ static void eval(const FusedArgs<real>& a, int i, int jj, int kk,
                                              bool owner, real& hx, real& hy, real& hz) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t nyz = a.nyz;
    const int64_t g = (int64_t)i * nyz + (int64_t)jj * nz + kk;
    const real K = -a.ch;
    const bool xin = i < nx - 1, yin = jj < ny - 1, zin = kk < nz - 1;
    const real* __restrict__ Ex = a.Ein[0];
    const real* __restrict__ Ey = a.Ein[1];
    const real* __restrict__ Ez = a.Ein[2];
    if (MASK & 1) {
      hx = a.Hin[0][g];
      if (yin && zin) {
        const int my = a.map_h[1][jj], mz = a.map_h[2][kk];
        const real dA = ddiff<real, EXACT>(Ez[g + nz], Ez[g], a.d[1], a.inv_d[1]);
        const real dB = ddiff<real, EXACT>(Ey[g + 1], Ey[g], a.d[2], a.inv_d[2]);
        real pa = 0, pb = 0, na = 0, nb = 0, ba = 0, ca = 0, bb = 0, cb = 0;
        if (my >= 0) { pa = a.psiHin[PSI_HXY - 6][pidx_y(i, my, kk, a.ph[1], nz)]; ba = a.b_h[1][jj]; ca = a.c_h[1][jj]; }
        if (mz >= 0) { pb = a.psiHin[PSI_HXZ - 6][pidx_z(i, jj, mz, ny, a.ph[2])]; bb = a.b_h[2][kk]; cb = a.c_h[2][kk]; }
        hx = curl_update(hx, K, true, dA, my >= 0, pa, ba, ca, &na, true, dB, mz >= 0, pb, bb, cb, &nb);
        if (owner) {
          if (my >= 0) a.psiHout[PSI_HXY - 6][pidx_y(i, my, kk, a.ph[1], nz)] = na;
          if (mz >= 0) a.psiHout[PSI_HXZ - 6][pidx_z(i, jj, mz, ny, a.ph[2])] = nb;
        }
      }
    }
    if (MASK & 2) {
      hy = a.Hin[1][g];
      if (zin && xin) {
        const int mz = a.map_h[2][kk], mx = a.map_h[0][i];
        const real dA = ddiff<real, EXACT>(Ex[g + 1], Ex[g], a.d[2], a.inv_d[2]);
        const real dB = ddiff<real, EXACT>(Ez[g + nyz], Ez[g], a.d[0], a.inv_d[0]);
        real pa = 0, pb = 0, na = 0, nb = 0, ba = 0, ca = 0, bb = 0, cb = 0;
        if (mz >= 0) { pa = a.psiHin[PSI_HYZ - 6][pidx_z(i, jj, mz, ny, a.ph[2])]; ba = a.b_h[2][kk]; ca = a.c_h[2][kk]; }
        if (mx >= 0) { pb = a.psiHin[PSI_HYX - 6][pidx_x(mx, jj, kk, ny, nz)]; bb = a.b_h[0][i]; cb = a.c_h[0][i]; }
        hy = curl_update(hy, K, true, dA, mz >= 0, pa, ba, ca, &na, true, dB, mx >= 0, pb, bb, cb, &nb);
        if (owner) {
          if (mz >= 0) a.psiHout[PSI_HYZ - 6][pidx_z(i, jj, mz, ny, a.ph[2])] = na;
          if (mx >= 0) a.psiHout[PSI_HYX - 6][pidx_x(mx, jj, kk, ny, nz)] = nb;
        }
      }
    }
    if (MASK & 4) {
      hz = a.Hin[2][g];
      if (xin && yin) {
        const int mx = a.map_h[0][i], my = a.map_h[1][jj];
        const real dA = ddiff<real, EXACT>(Ey[g + nyz], Ey[g], a.d[0], a.inv_d[0]);
        const real dB = ddiff<real, EXACT>(Ex[g + nz], Ex[g], a.d[1], a.inv_d[1]);
        real pa = 0, pb = 0, na = 0, nb = 0, ba = 0, ca = 0, bb = 0, cb = 0;
        if (mx >= 0) { pa = a.psiHin[PSI_HZX - 6][pidx_x(mx, jj, kk, ny, nz)]; ba = a.b_h[0][i]; ca = a.c_h[0][i]; }
        if (my >= 0) { pb = a.psiHin[PSI_HZY - 6][pidx_y(i, my, kk, a.ph[1], nz)]; bb = a.b_h[1][jj]; cb = a.c_h[1][jj]; }
        hz = curl_update(hz, K, true, dA, mx >= 0, pa, ba, ca, &na, true, dB, my >= 0, pb, bb, cb, &nb);
        if (owner) {
          if (mx >= 0) a.psiHout[PSI_HZX - 6][pidx_x(mx, jj, kk, ny, nz)] = na;
          if (my >= 0) a.psiHout[PSI_HZY - 6][pidx_y(i, my, kk, a.ph[1], nz)] = nb;
        }
      }
    }
    if (owner) {
      if (MASK & 1) a.Hout[0][g] = hx;
      if (MASK & 2) a.Hout[1][g] = hy;
      if (MASK & 4) a.Hout[2][g] = hz;
    }
  }
};

template <typename real, bool EXACT, int TY>
__global__ __launch_bounds__(TY * 64) void k_yee_fused_reg(FusedArgs<real> a) {
  const int nblk = a.tiles_z * a.tiles_y * a.chunks_x;
  const int per_xcd = a.grid_blocks / 8;
  const int bid = blockIdx.x;
  const int logical = (bid % 8) * per_xcd + bid / 8;
  if (logical >= nblk) return;
  // x chunks fastest: blocks of one XCD sweep neighbouring x chunks of a tile
  const int cxi = logical % a.chunks_x;
  const int tzi = (logical / a.chunks_x) % a.tiles_z;
  const int tyi = logical / (a.chunks_x * a.tiles_z);
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int j = tyi * TY + (threadIdx.x >> 6);
  const int k = tzi * 64 + (threadIdx.x & 63);
  if (j >= ny || k >= nz) return;
  const int i0 = cxi * a.xchunk;
  const int i1 = min(nx, i0 + a.xchunk);
  real hy_prev = 0, hz_prev = 0, dummy = 0;
  if (i0 > 0) HNode<real, EXACT>::template eval<6>(a, i0 - 1, j, k, false, dummy, hy_prev, hz_prev);
  const int my = a.map_e[1][j], mz = a.map_e[2][k];
  const real bey = my >= 0 ? a.b_e[1][j] : real(0), cey = my >= 0 ? a.c_e[1][j] : real(0);
  const real bez = mz >= 0 ? a.b_e[2][k] : real(0), cez = mz >= 0 ? a.c_e[2][k] : real(0);
  for (int i = i0; i < i1; ++i) {
    real hx, hy, hz, hx_jm = 0, hz_jm = 0, hx_km = 0, hy_km = 0, t = 0;
    HNode<real, EXACT>::template eval<7>(a, i, j, k, true, hx, hy, hz);
    if (j >= 1) HNode<real, EXACT>::template eval<5>(a, i, j - 1, k, false, hx_jm, t, hz_jm);
    if (k >= 1) HNode<real, EXACT>::template eval<3>(a, i, j, k - 1, false, hx_km, hy_km, t);
    const int64_t g = (int64_t)i * a.nyz + (int64_t)j * nz + k;
    const real K = a.ce[g];
    real ex = a.Ein[0][g], ey = a.Ein[1][g], ez = a.Ein[2][g];
    const bool xin = i >= 1, yin = j >= 1, zin = k >= 1;
    const int mx = a.map_e[0][i];
    if (yin && zin) {
      const real dA = ddiff<real, EXACT>(hz, hz_jm, a.d[1], a.inv_d[1]);
      const real dB = ddiff<real, EXACT>(hy, hy_km, a.d[2], a.inv_d[2]);
      real pa = 0, pb = 0, na = 0, nb = 0;
      if (my >= 0) pa = a.psiE[PSI_EXY][pidx_y(i, my, k, a.pe[1], nz)];
      if (mz >= 0) pb = a.psiE[PSI_EXZ][pidx_z(i, j, mz, ny, a.pe[2])];
      ex = curl_update(ex, K, true, dA, my >= 0, pa, bey, cey, &na, true, dB, mz >= 0, pb, bez, cez, &nb);
      if (my >= 0) a.psiE[PSI_EXY][pidx_y(i, my, k, a.pe[1], nz)] = na;
      if (mz >= 0) a.psiE[PSI_EXZ][pidx_z(i, j, mz, ny, a.pe[2])] = nb;
    }
    if (zin && xin) {
      const real dA = ddiff<real, EXACT>(hx, hx_km, a.d[2], a.inv_d[2]);
      const real dB = ddiff<real, EXACT>(hz, hz_prev, a.d[0], a.inv_d[0]);
      real pa = 0, pb = 0, na = 0, nb = 0, bx = 0, cx = 0;
      if (mz >= 0) pa = a.psiE[PSI_EYZ][pidx_z(i, j, mz, ny, a.pe[2])];
      if (mx >= 0) { pb = a.psiE[PSI_EYX][pidx_x(mx, j, k, ny, nz)]; bx = a.b_e[0][i]; cx = a.c_e[0][i]; }
      ey = curl_update(ey, K, true, dA, mz >= 0, pa, bez, cez, &na, true, dB, mx >= 0, pb, bx, cx, &nb);
      if (mz >= 0) a.psiE[PSI_EYZ][pidx_z(i, j, mz, ny, a.pe[2])] = na;
      if (mx >= 0) a.psiE[PSI_EYX][pidx_x(mx, j, k, ny, nz)] = nb;
    }
    if (xin && yin) {
      const real dA = ddiff<real, EXACT>(hy, hy_prev, a.d[0], a.inv_d[0]);
      const real dB = ddiff<real, EXACT>(hx, hx_jm, a.d[1], a.inv_d[1]);
      real pa = 0, pb = 0, na = 0, nb = 0, bx = 0, cx = 0;
      if (mx >= 0) { pa = a.psiE[PSI_EZX][pidx_x(mx, j, k, ny, nz)]; bx = a.b_e[0][i]; cx = a.c_e[0][i]; }
      if (my >= 0) pb = a.psiE[PSI_EZY][pidx_y(i, my, k, a.pe[1], nz)];
      ez = curl_update(ez, K, true, dA, mx >= 0, pa, bx, cx, &na, true, dB, my >= 0, pb, bey, cey, &nb);
      if (mx >= 0) a.psiE[PSI_EZX][pidx_x(mx, j, k, ny, nz)] = na;
      if (my >= 0) a.psiE[PSI_EZY][pidx_y(i, my, k, a.pe[1], nz)] = nb;
    }
    a.Eout[0][g] = ex;
    a.Eout[1][g] = ey;
    a.Eout[2][g] = ez;
    hy_prev = hy;
    hz_prev = hz;
  }
}

// March kernel ("march2"): every thread runs the same code; roles only
// decide what is stored. A workgroup is NW wavefronts; wave w handles row
// j0-1+w and lane l handles z node k0-1+l, so wave 0 and lane 0 form the low
// H apron, wave NW-1 and lane 63 only supply E(j+1) / E(k+1), and the
// (NW-2) x 62 interior is owned. Per plane: E^n(i) rows are exchanged through
// LDS (y) and cross-lane shuffles (z) for the H update, then H^{n+1/2}(i)
// likewise for the E update; H(i-1) stays in registers. Every global load of
// plane i+1 is issued before plane i is computed. CPML membership is
// arithmetic (two contiguous slabs per axis), so no load depends on another.
// Slabs, lds_barrier, ld and st come from yee_common.h.

// DIST: how many planes ahead the loads are issued (1 or 2).
// BATCH: the launch covers the tiles of several simulations; tile.pad0 names
// the simulation, whose arguments come from the device table arg.sims (read
// once per workgroup, before the plane loop).
template <typename real, bool EXACT, int NW, bool PML, int DIST, bool BATCH = false>
__global__ __launch_bounds__(NW * 64) void k_yee_march(MarchArgs<real> arg) {
  constexpr uint32_t B = sizeof(real);
  constexpr int XS = MARCH_MAX_XCHUNK + 1;
  // Row exchange, double-buffered by plane parity so one barrier per plane
  // suffices: plane i publishes H(i) and E^n(i+1) together.
  __shared__ real sEx[2][NW * 64], sEz[2][NW * 64], sHx[2][NW * 64], sHz[2][NW * 64];
  __shared__ real sX[4][XS];    // b_e, c_e, b_h, c_h along x for the chunk
  if ((int)blockIdx.x >= arg.ntiles) return;
  const MarchTile tile = reinterpret_cast<const MarchTile*>(arg.tiles)[blockIdx.x];
  if (tile.i1 <= tile.i0) return;
  const MarchArgs<real> a = BATCH ? arg.sims[tile.pad0] : arg;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nyz = (size_t)ny * nz;
  // Field and ce arrays use rows of a.zrow slots (padded rows when a.zpad:
  // node k in slot zslot(k), each z tile of a.zpad nodes owning one
  // aligned 64-slot block); psi arrays keep dense rows of nz.
  const size_t fplane = (size_t)ny * a.zrow, fcell = (size_t)nx * fplane;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int jn = tile.jn, kn = tile.kn;
  const int j = tile.j0 - 1 + w, k = tile.k0 - 1 + l;
  const bool valid = j >= 0 && j < ny && k >= 0 && k < nz && w <= jn + 1 && l <= kn + 1;
  // roles: H is needed on the apron row/lane and the owned nodes; ce and the
  // E-side psi only on owned nodes; row jn+1 / lane kn+1 only supply E.
  const bool hrole = valid && w <= jn && l <= kn;
  const bool own = hrole && w >= 1 && l >= 1;
  const int jc = min(max(j, 0), ny - 1), kc = min(max(k, 0), nz - 1);
  const int i0 = tile.i0, i1 = tile.i1;
  const int ib = i0 > 0 ? i0 - 1 : 0;

  // derived bases
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
  // slot bases: E: ExY ExZ EyZ EyX EzX EzY ; H: HxY HxZ HyZ HyX HzX HzY
  real* pExY = a.psiE;
  real* pExZ = pExY + sy_e;
  real* pEyZ = pExZ + sz_e;
  real* pEyX = pEyZ + sz_e;
  real* pEzX = pEyX + sx_e;
  real* pEzY = pEzX + sx_e;
  const size_t hoff[6] = {0, sy_h, sy_h + sz_h, sy_h + 2 * sz_h, sy_h + 2 * sz_h + sx_h,
                          sy_h + 2 * sz_h + 2 * sx_h};
  const real* coef_y = a.coef + 4 * (size_t)nx;
  const real* coef_z = a.coef + 4 * (size_t)(nx + ny);

  // x coefficients of the chunk -> LDS (read after the first barrier)
  for (int t = tid; PML && t < 4 * (i1 - ib); t += NW * 64) {
    const int q = t / (i1 - ib), ii = ib + t % (i1 - ib);
    sX[q][ii - ib] = a.coef[(size_t)q * nx + ii];
  }

  const Slabs<real> ye{a.slab_e[1][0], a.slab_e[1][1]}, yh{a.slab_h[1][0], a.slab_h[1][1]};
  const Slabs<real> ze{a.slab_e[2][0], a.slab_e[2][1]}, zh{a.slab_h[2][0], a.slab_h[2][1]};
  const Slabs<real> xe{a.slab_e[0][0], a.slab_e[0][1]}, xh{a.slab_h[0][0], a.slab_h[0][1]};
  const bool pyh = PML && hrole && yh.in(jc), pzh = PML && hrole && zh.in(kc);
  const bool pye = PML && own && ye.in(jc), pze = PML && own && ze.in(kc);
  const real bye = pye ? coef_y[jc] : real(0), cye = pye ? coef_y[ny + jc] : real(0);
  const real byh = pyh ? coef_y[2 * ny + jc] : real(0), cyh = pyh ? coef_y[3 * ny + jc] : real(0);
  const real bze = pze ? coef_z[kc] : real(0), cze = pze ? coef_z[nz + kc] : real(0);
  const real bzh = pzh ? coef_z[2 * nz + kc] : real(0), czh = pzh ? coef_z[3 * nz + kc] : real(0);
  // byte offsets inside one plane of each layout
  const uint32_t col = (uint32_t)(jc * nz + kc) * B;   // psi (dense rows)
  // fields and ce: load slot of this lane's node; store slot: the owned
  // block of the tile (lane l >= 1 owns node slot l - 1; lane 0 and the
  // lanes past the block's nodes write its pad slots, so a padded row store
  // covers whole lines)
  const uint32_t colf = (uint32_t)(jc * a.zrow + zslot(kc, a.zpad)) * B;
  const uint32_t cols =
      a.zpad ? (uint32_t)(jc * a.zrow + tile.k0 / a.zpad * ZPAD_BLOCK +
                          (l == 0 ? a.zpad : l <= a.zpad ? l - 1 : l)) * B
             : colf;
  // lanes that store fields: the owned nodes; with padded rows every lane of
  // an owned row (its slot is an owned node, a pad slot or past the row end)
  const bool st_lane = a.zpad ? (j >= 0 && j < ny && w >= 1 && w <= jn) : own;
  const uint32_t cyh_o = (uint32_t)(yh.m(jc) * nz + kc) * B;
  const uint32_t czh_o = (uint32_t)(jc * a.ph[2] + zh.m(kc)) * B;
  const uint32_t cye_o = (uint32_t)(ye.m(jc) * nz + kc) * B;
  const uint32_t cze_o = (uint32_t)(jc * a.pe[2] + ze.m(kc)) * B;
  const size_t ys_h = (size_t)a.ph[1] * nz, zs_h = (size_t)ny * a.ph[2];
  const size_t ys_e = (size_t)a.pe[1] * nz, zs_e = (size_t)ny * a.pe[2];

  // Everything plane i needs from memory, fetched one plane ahead in one
  // group so no later load forces a wait on it.
  // CPML psi of plane i (H side: HxY HxZ HyZ HyX HzX HzY; E side: ExY ExZ EyZ
  // EyX EzX EzY), fetched with the rest of the plane.
  struct Psi {
    real ph[6], pe[6];
  };
  struct Plane {
    real e[3];    // E^n(i+1) at this node
    real h[3];    // H^{n-1/2}(i)
    real ce;
    Psi q;
  };

  auto fetch_psi = [&](Psi& Q, int i) {
    for (int q = 0; q < 6; ++q) Q.ph[q] = Q.pe[q] = real(0);
    if (!PML) return;
    if (pyh) {
      Q.ph[0] = ld(a.psiHin + hoff[0] + i * ys_h, cyh_o);
      Q.ph[5] = ld(a.psiHin + hoff[5] + i * ys_h, cyh_o);
    }
    if (pzh) {
      Q.ph[1] = ld(a.psiHin + hoff[1] + i * zs_h, czh_o);
      Q.ph[2] = ld(a.psiHin + hoff[2] + i * zs_h, czh_o);
    }
    if (hrole && xh.in(i)) {
      const size_t xo = (size_t)xh.m(i) * nyz;
      Q.ph[3] = ld(a.psiHin + hoff[3] + xo, col);
      Q.ph[4] = ld(a.psiHin + hoff[4] + xo, col);
    }
    if (pye) {
      Q.pe[0] = ld(pExY + i * ys_e, cye_o);
      Q.pe[5] = ld(pEzY + i * ys_e, cye_o);
    }
    if (pze) {
      Q.pe[1] = ld(pExZ + i * zs_e, cze_o);
      Q.pe[2] = ld(pEyZ + i * zs_e, cze_o);
    }
    if (own && xe.in(i)) {
      const size_t xo = (size_t)xe.m(i) * nyz;
      Q.pe[3] = ld(pEyX + xo, col);
      Q.pe[4] = ld(pEzX + xo, col);
    }
  };  auto fetch = [&](Plane& P, int i) {
    const size_t o = (size_t)i * fplane;
    if (valid && i + 1 < nx) {
      const size_t on = o + fplane;
      P.e[0] = ld(Ein0 + on, colf);
      P.e[1] = ld(Ein1 + on, colf);
      P.e[2] = ld(Ein2 + on, colf);
    } else {
      P.e[0] = P.e[1] = P.e[2] = real(0);
    }
    P.h[0] = P.h[1] = P.h[2] = P.ce = real(0);
    if (hrole) {
      P.h[0] = ld(Hin0 + o, colf);
      P.h[1] = ld(Hin1 + o, colf);
      P.h[2] = ld(Hin2 + o, colf);
    }
    if (own) P.ce = ld(a.ce + o, colf);
    fetch_psi(P.q, i);
  };

  real ecur[3] = {0, 0, 0};
  if (valid) {
    const size_t o = (size_t)ib * fplane;
    ecur[0] = ld(Ein0 + o, colf);
    ecur[1] = ld(Ein1 + o, colf);
    ecur[2] = ld(Ein2 + o, colf);
  }
  real hy_prev = 0, hz_prev = 0;
  const real K_h = -a.ch;
  // One plane: P holds plane i's memory operands; E(i) is in ecur.
  auto plane = [&](int i, const Plane& P, Plane& Next) {
    if (i + DIST < i1) fetch(Next, i + DIST);
    const Psi& Q = P.q;
    const bool live = i >= i0;                // the plane before i0 only rebuilds H(i0-1)
    const int bi = i & 1;
    const real ex_jp = w < NW - 1 ? sEx[bi][min(tid + 64, NW * 64 - 1)] : real(0);
    const real ez_jp = w < NW - 1 ? sEz[bi][min(tid + 64, NW * 64 - 1)] : real(0);
    const real ex_kp = __shfl_down(ecur[0], 1);
    const real ey_kp = __shfl_down(ecur[1], 1);
    // ---- H^{n+1/2}(i) ----
    const bool xh_in = i < nx - 1, yh_in = j < ny - 1, zh_in = k < nz - 1;
    const bool pxh = PML && xh.in(i);
    const real bxh = sX[2][i - ib], cxh = sX[3][i - ib];
    real hx = P.h[0], hy = P.h[1], hz = P.h[2];
    real nps[6] = {0, 0, 0, 0, 0, 0};
    if (hrole) {
      if (yh_in && zh_in) {
        const real dA = ddiff<real, EXACT>(ez_jp, ecur[2], a.d[1], a.inv_d[1]);
        const real dB = ddiff<real, EXACT>(ey_kp, ecur[1], a.d[2], a.inv_d[2]);
        hx = curl_update(hx, K_h, true, dA, pyh, Q.ph[0], byh, cyh, &nps[0], true, dB, pzh,
                         Q.ph[1], bzh, czh, &nps[1]);
      }
      if (zh_in && xh_in) {
        const real dA = ddiff<real, EXACT>(ex_kp, ecur[0], a.d[2], a.inv_d[2]);
        const real dB = ddiff<real, EXACT>(P.e[2], ecur[2], a.d[0], a.inv_d[0]);
        hy = curl_update(hy, K_h, true, dA, pzh, Q.ph[2], bzh, czh, &nps[2], true, dB, pxh,
                         Q.ph[3], bxh, cxh, &nps[3]);
      }
      if (xh_in && yh_in) {
        const real dA = ddiff<real, EXACT>(P.e[1], ecur[1], a.d[0], a.inv_d[0]);
        const real dB = ddiff<real, EXACT>(ex_jp, ecur[0], a.d[1], a.inv_d[1]);
        hz = curl_update(hz, K_h, true, dA, pxh, Q.ph[4], bxh, cxh, &nps[4], true, dB, pyh,
                         Q.ph[5], byh, cyh, &nps[5]);
      }
      if (own && live) {
        if (pyh && yh_in && zh_in) st(a.psiHout + hoff[0] + i * ys_h, cyh_o, nps[0]);
        if (pzh && yh_in && zh_in) st(a.psiHout + hoff[1] + i * zs_h, czh_o, nps[1]);
        if (pzh && zh_in && xh_in) st(a.psiHout + hoff[2] + i * zs_h, czh_o, nps[2]);
        if (pxh && zh_in && xh_in) st(a.psiHout + hoff[3] + (size_t)xh.m(i) * nyz, col, nps[3]);
        if (pxh && xh_in && yh_in) st(a.psiHout + hoff[4] + (size_t)xh.m(i) * nyz, col, nps[4]);
        if (pyh && xh_in && yh_in) st(a.psiHout + hoff[5] + i * ys_h, cyh_o, nps[5]);
      }
    }
    // Field rows are stored with the streaming hint: the next read of a row
    // is the next step's, long after it has left the caches, and the lines
    // it would push out hold the y-apron rows the neighbouring tiles read.
    if (st_lane && live) {
      const size_t o = (size_t)i * fplane;
      st_nt(Ho0 + o, cols, hx);
      st_nt(Ho1 + o, cols, hy);
      st_nt(Ho2 + o, cols, hz);
    }
    sHx[bi][tid] = hx;
    sHz[bi][tid] = hz;
    sEx[bi ^ 1][tid] = P.e[0];    // E^n(i+1) for the H update of plane i+1
    sEz[bi ^ 1][tid] = P.e[2];
    lds_barrier();
    const real hx_jm = w > 0 ? sHx[bi][tid - 64] : real(0);
    const real hz_jm = w > 0 ? sHz[bi][tid - 64] : real(0);
    const real hx_km = __shfl_up(hx, 1);
    const real hy_km = __shfl_up(hy, 1);
    // ---- E^{n+1}(i) ----
    if (st_lane && live) {
      const bool xe_in = i >= 1, ye_in = j >= 1, ze_in = k >= 1;
      const bool pxe = PML && xe.in(i);
      const real bxe = sX[0][i - ib], cxe = sX[1][i - ib];
      const real K = P.ce;
      real ex = ecur[0], ey = ecur[1], ez = ecur[2];
      real npe[6] = {0, 0, 0, 0, 0, 0};
      if (ye_in && ze_in) {
        const real dA = ddiff<real, EXACT>(hz, hz_jm, a.d[1], a.inv_d[1]);
        const real dB = ddiff<real, EXACT>(hy, hy_km, a.d[2], a.inv_d[2]);
        ex = curl_update(ex, K, true, dA, pye, Q.pe[0], bye, cye, &npe[0], true, dB, pze,
                         Q.pe[1], bze, cze, &npe[1]);
      }
      if (ze_in && xe_in) {
        const real dA = ddiff<real, EXACT>(hx, hx_km, a.d[2], a.inv_d[2]);
        const real dB = ddiff<real, EXACT>(hz, hz_prev, a.d[0], a.inv_d[0]);
        ey = curl_update(ey, K, true, dA, pze, Q.pe[2], bze, cze, &npe[2], true, dB, pxe,
                         Q.pe[3], bxe, cxe, &npe[3]);
      }
      if (xe_in && ye_in) {
        const real dA = ddiff<real, EXACT>(hy, hy_prev, a.d[0], a.inv_d[0]);
        const real dB = ddiff<real, EXACT>(hx, hx_jm, a.d[1], a.inv_d[1]);
        ez = curl_update(ez, K, true, dA, pxe, Q.pe[4], bxe, cxe, &npe[4], true, dB, pye,
                         Q.pe[5], bye, cye, &npe[5]);
      }
      const size_t o = (size_t)i * fplane;
      st_nt(Eo0 + o, cols, ex);
      st_nt(Eo1 + o, cols, ey);
      st_nt(Eo2 + o, cols, ez);
      if (pye && ye_in && ze_in) st(pExY + i * ys_e, cye_o, npe[0]);
      if (pze && ye_in && ze_in) st(pExZ + i * zs_e, cze_o, npe[1]);
      if (pze && ze_in && xe_in) st(pEyZ + i * zs_e, cze_o, npe[2]);
      if (own && pxe && ze_in && xe_in) st(pEyX + (size_t)xe.m(i) * nyz, col, npe[3]);
      if (own && pxe && xe_in && ye_in) st(pEzX + (size_t)xe.m(i) * nyz, col, npe[4]);
      if (pye && xe_in && ye_in) st(pEzY + i * ys_e, cye_o, npe[5]);
    }
    hy_prev = hy;
    hz_prev = hz;
  };
  sEx[ib & 1][tid] = ecur[0];
  sEz[ib & 1][tid] = ecur[2];
  lds_barrier();
  if constexpr (DIST == 1) {
    // Two plane buffers alternate so no register still being loaded is copied.
    Plane A, Bq;
    fetch(A, ib);
    for (int i = ib; i < i1; i += 2) {
      if (i > ib) {
        ecur[0] = Bq.e[0];
        ecur[1] = Bq.e[1];
        ecur[2] = Bq.e[2];
      }
      plane(i, A, Bq);
      if (i + 1 >= i1) break;
      ecur[0] = A.e[0];
      ecur[1] = A.e[1];
      ecur[2] = A.e[2];
      plane(i + 1, Bq, A);
    }
  } else {
    // Three plane buffers rotate: plane i is computed from one while planes
    // i + 1 and i + 2 are in flight in the others. E^n(i) is copied out of
    // the buffer of plane i - 1 before that buffer receives plane i + 2.
    Plane A, Bq, Cq;
    fetch(A, ib);
    if (ib + 1 < i1) fetch(Bq, ib + 1);
    for (int i = ib; i < i1; i += 3) {
      if (i > ib) {
        ecur[0] = Cq.e[0];
        ecur[1] = Cq.e[1];
        ecur[2] = Cq.e[2];
      }
      plane(i, A, Cq);
      if (i + 1 >= i1) break;
      ecur[0] = A.e[0];
      ecur[1] = A.e[1];
      ecur[2] = A.e[2];
      plane(i + 1, Bq, A);
      if (i + 2 >= i1) break;
      ecur[0] = Bq.e[0];
      ecur[1] = Bq.e[1];
      ecur[2] = Bq.e[2];
      plane(i + 2, Cq, Bq);
    }
  }
}

// ADE for the ping-pong layout: E^n comes from the input level, the
// curl-updated E^{n+1} is corrected in place (simulation.py:638-669).
template <typename real>
__global__ void k_ade_pp(const real* __restrict__ Ein0, const real* __restrict__ Ein1,
                         const real* __restrict__ Ein2, real* __restrict__ Eout0,
                         real* __restrict__ Eout1, real* __restrict__ Eout2,
                         const int64_t* __restrict__ cells, int64_t n, real* __restrict__ Pcur,
                         real* __restrict__ Pprev, const real* __restrict__ abc, int n_poles,
                         real inv_eps) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || n_poles == 0) return;
  const real* Ein[3] = {Ein0, Ein1, Ein2};
  real* Eout[3] = {Eout0, Eout1, Eout2};
  ade_pp_node<real>(Ein, Eout, cells, n, Pcur, Pprev, abc, n_poles, inv_eps, t);
}

// Batched ADE: job blockIdx.y, node blockIdx.x * blockDim.x + lane.
template <typename real>
__global__ void k_ade_jobs(const AdeJob<real>* __restrict__ jobs) {
  const AdeJob<real>& J = jobs[blockIdx.y];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= J.n || J.n_poles == 0) return;
  ade_pp_node<real>(J.Ein, J.Eout, J.cells, J.n, J.Pcur, J.Pprev, J.abc, J.n_poles, J.inv_eps, t);
}

template <typename real>
void launch_ade_jobs(const AdeJob<real>* jobs, int n_jobs, int64_t max_n, hipStream_t s) {
  if (n_jobs <= 0 || max_n <= 0) return;
  dim3 grid((unsigned)((max_n + 255) / 256), (unsigned)n_jobs);
  k_ade_jobs<real><<<grid, 256, 0, s>>>(jobs);
}
template void launch_ade_jobs<float>(const AdeJob<float>*, int, int64_t, hipStream_t);
template void launch_ade_jobs<double>(const AdeJob<double>*, int, int64_t, hipStream_t);


template <typename real>
bool fused_supported(const StepArgs<real>& a) {
  return a.nx >= 2 && a.ny >= 2 && a.nz >= 2;
}

template <typename real>
void launch_fused(FusedArgs<real>& a, bool exact, hipStream_t s) {
  constexpr int TY = 4;
  a.tiles_z = (a.nz + TZ - 1) / TZ;
  a.tiles_y = (a.ny + TY - 1) / TY;
  a.chunks_x = (a.nx + a.xchunk - 1) / a.xchunk;
  const int nblk = a.tiles_z * a.tiles_y * a.chunks_x;
  a.grid_blocks = (nblk + 7) / 8 * 8;
  if (a.variant == 1) {
    if (exact) k_yee_fused<real, true, TY><<<a.grid_blocks, TY * 64, 0, s>>>(a);
    else k_yee_fused<real, false, TY><<<a.grid_blocks, TY * 64, 0, s>>>(a);
  } else {
    if (exact) k_yee_fused_reg<real, true, TY><<<a.grid_blocks, TY * 64, 0, s>>>(a);
    else k_yee_fused_reg<real, false, TY><<<a.grid_blocks, TY * 64, 0, s>>>(a);
  }
}

template <typename real>
void launch_ade_pp(real* const* Ein, real* const* Eout, const int64_t* cells, int64_t n, real* Pcur,
                   real* Pprev, const real* abc, int n_poles, real inv_eps, hipStream_t s) {
  k_ade_pp<real><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(Ein[0], Ein[1], Ein[2], Eout[0],
                                                             Eout[1], Eout[2], cells, n, Pcur,
                                                             Pprev, abc, n_poles, inv_eps);
}

template <typename real, int NW, int DIST>
void launch_march_nw(MarchArgs<real>& a, bool exact, bool pml, hipStream_t s) {
  if (a.sims) {   // batch
    if (pml) {
      if (exact) k_yee_march<real, true, NW, true, DIST, true><<<a.ntiles, NW * 64, 0, s>>>(a);
      else k_yee_march<real, false, NW, true, DIST, true><<<a.ntiles, NW * 64, 0, s>>>(a);
    } else {
      if (exact) k_yee_march<real, true, NW, false, DIST, true><<<a.ntiles, NW * 64, 0, s>>>(a);
      else k_yee_march<real, false, NW, false, DIST, true><<<a.ntiles, NW * 64, 0, s>>>(a);
    }
    return;
  }
  if (pml) {
    if (exact) k_yee_march<real, true, NW, true, DIST><<<a.ntiles, NW * 64, 0, s>>>(a);
    else k_yee_march<real, false, NW, true, DIST><<<a.ntiles, NW * 64, 0, s>>>(a);
  } else {
    if (exact) k_yee_march<real, true, NW, false, DIST><<<a.ntiles, NW * 64, 0, s>>>(a);
    else k_yee_march<real, false, NW, false, DIST><<<a.ntiles, NW * 64, 0, s>>>(a);
  }
}

template <typename real>
void launch_march(MarchArgs<real>& a, bool exact, bool pml, int nw, int dist, hipStream_t s) {
  if (a.ntiles == 0) return;
  if (nw == 16) {
    if (dist == 2) launch_march_nw<real, 16, 2>(a, exact, pml, s);
    else launch_march_nw<real, 16, 1>(a, exact, pml, s);
  } else if (nw == 12) {
    if (dist == 2) launch_march_nw<real, 12, 2>(a, exact, pml, s);
    else launch_march_nw<real, 12, 1>(a, exact, pml, s);
  } else {
    if (dist == 2) launch_march_nw<real, MARCH_NW, 2>(a, exact, pml, s);
    else launch_march_nw<real, MARCH_NW, 1>(a, exact, pml, s);
  }
}

template void launch_march<float>(MarchArgs<float>&, bool, bool, int, int, hipStream_t);
template void launch_march<double>(MarchArgs<double>&, bool, bool, int, int, hipStream_t);
template bool fused_supported<float>(const StepArgs<float>&);
template bool fused_supported<double>(const StepArgs<double>&);
template void launch_fused<float>(FusedArgs<float>&, bool, hipStream_t);
template void launch_fused<double>(FusedArgs<double>&, bool, hipStream_t);
template void launch_ade_pp<float>(float* const*, float* const*, const int64_t*, int64_t, float*,
                                   float*, const float*, int, float, hipStream_t);
template void launch_ade_pp<double>(double* const*, double* const*, const int64_t*, int64_t,
                                    double*, double*, const double*, int, double, hipStream_t);

}  // namespace pfhip
