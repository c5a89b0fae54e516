// CPU FDTD stepper: a C++/OpenMP restatement of the reference's compiled CPU
// backend, used as the measured CPU baseline and as a second CPU check.
//
// TEST AND BASELINE INFRASTRUCTURE ONLY (like everything under oracle/): the
// benchmark's cpu_baseline leg and the tests load it; the product never does.
//
// Follows rust/src/kernel.rs:153-372 (fused H pass Hx+Hy+Hz per node, then the
// fused E pass; differences multiplied by 1/d; the CPML psi update
// psi = b psi + c d folded into the curl term, F += K (cA - cB) with
// cX = d + psi where the node has psi) and rust/src/blocked.rs:134-350
// (ghost-zone temporal blocking: x-slab tiles evolve T steps in a private
// buffer holding the slab plus a halo of T + 2 rows, halo rows copied from
// snapshots taken before any tile writes back, only the core written back;
// bit-identical to plain stepping). Sources are added after their pass
// (H at (n+1/2) dt, E at (n+1) dt; rustbackend.py:151-165). Beyond the
// reference's CPU backend, the plain path also steps the ADE polarisation of
// dispersive media (simulation.py:638-669) and accumulates DFT monitors every
// sampled step in complex128 (simulation.py:1116-1133), so it runs the same
// workload as the GPU stepper. Rows are shared out with OpenMP.
//
// Arithmetic: no contraction (-ffp-contract=off), coefficients rounded to the
// field type as the reference casts its tables (rustbackend.py:140-146).
#include <omp.h>

#include <algorithm>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/pfhip.h"

namespace {

thread_local std::string g_err;

enum Psi { EXY, EXZ, EYZ, EYX, EZX, EZY, HXY, HXZ, HYZ, HYX, HZX, HZY };

// Pointers and extents of one domain or one tile buffer.
template <typename T>
struct View {
  T* f[6];
  T* psi[12];
  const T* be[3];
  const T* cE[3];
  const T* bh[3];
  const T* cH[3];
  const int* me[3];
  const int* mh[3];
  int pe[3], ph[3];        // compact psi extents (x extents used only for sizes)
  const T* ce;
  T ch;
  T inv[3];
  int nx, ny, nz;

  size_t at(int i, int j, int k) const { return ((size_t)i * ny + j) * nz + k; }

  // One psi-corrected difference: d, plus the updated psi when the node has one.
  static T term(T d, T* ps, size_t pid, T b, T c, bool has) {
    if (!has) return d;
    const T p = b * ps[pid] + c * d;
    ps[pid] = p;
    return d + p;
  }

  void h_rows(int i0, int i1) const {
    const size_t sx = (size_t)ny * nz, sy = nz;
    const int fx = nx > 1 ? nx - 1 : 1, fy = ny > 1 ? ny - 1 : 1, fz = nz > 1 ? nz - 1 : 1;
    T *Ex = f[0], *Ey = f[1], *Ez = f[2], *Hx = f[3], *Hy = f[4], *Hz = f[5];
    for (int i = i0; i < i1; ++i) {
      const int mi = nx > 1 ? mh[0][i] : -1;
      for (int j = 0; j < ny; ++j) {
        const int mj = ny > 1 ? mh[1][j] : -1;
        for (int k = 0; k < nz; ++k) {
          const size_t id = at(i, j, k);
          const int mk = nz > 1 ? mh[2][k] : -1;
          if (j < fy && k < fz) {   // Hx
            T cy = 0, cz = 0;
            if (ny > 1)
              cy = term((Ez[id + sy] - Ez[id]) * inv[1], psi[HXY], ((size_t)i * ph[1] + mj) * nz + k,
                        bh[1][j], cH[1][j], mj >= 0);
            if (nz > 1)
              cz = term((Ey[id + 1] - Ey[id]) * inv[2], psi[HXZ], ((size_t)i * ny + j) * ph[2] + mk,
                        bh[2][k], cH[2][k], mk >= 0);
            Hx[id] = Hx[id] + (-ch) * (cy - cz);
          }
          if (i < fx && k < fz) {   // Hy
            T cz = 0, cx = 0;
            if (nz > 1)
              cz = term((Ex[id + 1] - Ex[id]) * inv[2], psi[HYZ], ((size_t)i * ny + j) * ph[2] + mk,
                        bh[2][k], cH[2][k], mk >= 0);
            if (nx > 1)
              cx = term((Ez[id + sx] - Ez[id]) * inv[0], psi[HYX], ((size_t)mi * ny + j) * nz + k,
                        bh[0][i], cH[0][i], mi >= 0);
            Hy[id] = Hy[id] + (-ch) * (cz - cx);
          }
          if (i < fx && j < fy) {   // Hz
            T cx = 0, cy = 0;
            if (nx > 1)
              cx = term((Ey[id + sx] - Ey[id]) * inv[0], psi[HZX], ((size_t)mi * ny + j) * nz + k,
                        bh[0][i], cH[0][i], mi >= 0);
            if (ny > 1)
              cy = term((Ex[id + sy] - Ex[id]) * inv[1], psi[HZY], ((size_t)i * ph[1] + mj) * nz + k,
                        bh[1][j], cH[1][j], mj >= 0);
            Hz[id] = Hz[id] + (-ch) * (cx - cy);
          }
        }
      }
    }
  }

  void e_rows(int i0, int i1) const {
    const size_t sx = (size_t)ny * nz, sy = nz;
    const int bx = nx > 1, by = ny > 1, bz = nz > 1;
    T *Ex = f[0], *Ey = f[1], *Ez = f[2], *Hx = f[3], *Hy = f[4], *Hz = f[5];
    for (int i = i0; i < i1; ++i) {
      const int mi = nx > 1 ? me[0][i] : -1;
      for (int j = 0; j < ny; ++j) {
        const int mj = ny > 1 ? me[1][j] : -1;
        for (int k = 0; k < nz; ++k) {
          const size_t id = at(i, j, k);
          const int mk = nz > 1 ? me[2][k] : -1;
          const T K = ce[id];
          if (j >= by && k >= bz) {   // Ex
            T cy = 0, cz = 0;
            if (ny > 1)
              cy = term((Hz[id] - Hz[id - sy]) * inv[1], psi[EXY], ((size_t)i * pe[1] + mj) * nz + k,
                        be[1][j], cE[1][j], mj >= 0);
            if (nz > 1)
              cz = term((Hy[id] - Hy[id - 1]) * inv[2], psi[EXZ], ((size_t)i * ny + j) * pe[2] + mk,
                        be[2][k], cE[2][k], mk >= 0);
            Ex[id] = Ex[id] + K * (cy - cz);
          }
          if (i >= bx && k >= bz) {   // Ey
            T cz = 0, cx = 0;
            if (nz > 1)
              cz = term((Hx[id] - Hx[id - 1]) * inv[2], psi[EYZ], ((size_t)i * ny + j) * pe[2] + mk,
                        be[2][k], cE[2][k], mk >= 0);
            if (nx > 1)
              cx = term((Hz[id] - Hz[id - sx]) * inv[0], psi[EYX], ((size_t)mi * ny + j) * nz + k,
                        be[0][i], cE[0][i], mi >= 0);
            Ey[id] = Ey[id] + K * (cz - cx);
          }
          if (i >= bx && j >= by) {   // Ez
            T cx = 0, cy = 0;
            if (nx > 1)
              cx = term((Hy[id] - Hy[id - sx]) * inv[0], psi[EZX], ((size_t)mi * ny + j) * nz + k,
                        be[0][i], cE[0][i], mi >= 0);
            if (ny > 1)
              cy = term((Hx[id] - Hx[id - sy]) * inv[1], psi[EZY], ((size_t)i * pe[1] + mj) * nz + k,
                        be[1][j], cE[1][j], mj >= 0);
            Ez[id] = Ez[id] + K * (cx - cy);
          }
        }
      }
    }
  }
};

// Rows per parallel chunk: a few chunks per thread.
template <typename F>
void for_rows(int n, F&& fn) {
  const int threads = omp_get_max_threads();
  const int chunk = std::max(1, n / (threads * 4));
  const int chunks = (n + chunk - 1) / chunk;
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < chunks; ++c) fn(c * chunk, std::min(n, (c + 1) * chunk));
}

struct Source {
  int comp;
  int i, j, k;
  double amp;
  int row;
};

template <typename T>
struct Stepper {
  int nx, ny, nz;
  int64_t n_steps;
  std::vector<T> f[6], ce, psi[12], coef[12];   // coef: be, cE, bh, cH per axis
  std::vector<int> me[3], mh[3];
  int pe[3] = {0, 0, 0}, ph[3] = {0, 0, 0};
  T ch, inv[3];
  std::vector<Source> src;
  std::vector<double> wave;   // [n_wave][n_steps]
  struct Ade {
    std::vector<int64_t> cells;
    int np;
    std::vector<T> a, b, c;
    T inv_eps;
    std::vector<T> P, Pp, En;   // [pole][comp][cell], En [comp][cell]
  };
  std::vector<Ade> ade;
  struct Dft {
    int n_comp, comps[6], n_freq, interval;
    pf_box box;
    int64_t nbox, n_samples;
    std::vector<double> ph_e, ph_h;   // [sample][freq][2]
    double dt;
    std::vector<std::complex<double>> acc;   // [comp][freq][node]
  };
  std::vector<Dft> dft;
  int t_block = 0, tile_rows = 0;

  explicit Stepper(const pf_desc& d) {
    nx = (int)d.shape[0];
    ny = (int)d.shape[1];
    nz = (int)d.shape[2];
    const size_t n = (size_t)nx * ny * nz;
    for (auto& v : f) v.assign(n, T(0));
    const T* ce_in = static_cast<const T*>(d.ce);
    ce.assign(ce_in, ce_in + n);
    ch = (T)d.ch;
    const int na[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
      inv[a] = (T)(1.0 / d.cell[a]);
      const double* src_c[4] = {d.b_e[a], d.c_e[a], d.b_h[a], d.c_h[a]};
      for (int q = 0; q < 4; ++q) {
        coef[4 * a + q].resize(na[a]);
        for (int x = 0; x < na[a]; ++x) coef[4 * a + q][x] = (T)src_c[q][x];
      }
      // compact psi index of the nodes whose (field-type) c is non-zero
      me[a].assign(na[a], -1);
      mh[a].assign(na[a], -1);
      for (int x = 0; x < na[a]; ++x) {
        if (coef[4 * a + 1][x] != T(0)) me[a][x] = pe[a]++;
        if (coef[4 * a + 3][x] != T(0)) mh[a][x] = ph[a]++;
      }
    }
    const size_t ex_y = (size_t)nx * pe[1] * nz, ex_z = (size_t)nx * ny * pe[2],
                 ex_x = (size_t)pe[0] * ny * nz;
    const size_t hx_y = (size_t)nx * ph[1] * nz, hx_z = (size_t)nx * ny * ph[2],
                 hx_x = (size_t)ph[0] * ny * nz;
    const size_t sz[12] = {ex_y, ex_z, ex_z, ex_x, ex_x, ex_y, hx_y, hx_z, hx_z, hx_x, hx_x, hx_y};
    for (int p = 0; p < 12; ++p) psi[p].assign(std::max<size_t>(sz[p], 1), T(0));
    n_steps = d.n_steps;
    wave.assign(d.wave, d.wave + (size_t)d.n_wave * d.n_steps);
    for (int s = 0; s < d.n_src; ++s) {
      const int64_t c = d.src_cell[s];
      src.push_back({d.src_comp[s], (int)(c / ((int64_t)ny * nz)), (int)(c / nz % ny), (int)(c % nz),
                     d.src_amp[s], d.src_wave[s]});
    }
    for (int m = 0; m < d.n_ade; ++m) {
      const pf_ade_desc& a = d.ade[m];
      Ade A;
      A.cells.assign(a.cells, a.cells + a.n_cells);
      A.np = a.n_poles;
      for (int p = 0; p < a.n_poles; ++p) {
        A.a.push_back((T)a.abc[3 * p]);
        A.b.push_back((T)a.abc[3 * p + 1]);
        A.c.push_back((T)a.abc[3 * p + 2]);
      }
      A.inv_eps = (T)a.inv_eps;
      A.P.assign((size_t)a.n_poles * 3 * a.n_cells, T(0));
      A.Pp = A.P;
      A.En.assign(3 * (size_t)a.n_cells, T(0));
      ade.push_back(std::move(A));
    }
    for (int m = 0; m < d.n_dft; ++m) {
      const pf_dft_desc& s = d.dft[m];
      Dft D;
      D.n_comp = s.n_comp;
      std::copy(s.comps, s.comps + 6, D.comps);
      D.n_freq = s.n_freq;
      D.interval = s.interval;
      D.box = s.box;
      D.nbox = s.box.count[0] * s.box.count[1] * s.box.count[2];
      D.n_samples = s.n_samples;
      D.ph_e.assign(s.phase_e, s.phase_e + 2 * (size_t)s.n_samples * s.n_freq);
      D.ph_h.assign(s.phase_h, s.phase_h + 2 * (size_t)s.n_samples * s.n_freq);
      D.dt = s.dt;
      D.acc.assign((size_t)s.n_comp * s.n_freq * D.nbox, std::complex<double>(0, 0));
      dft.push_back(std::move(D));
    }
    if (d.n_flux > 0) throw std::invalid_argument("flux monitors are not restated on the CPU");
    if (d.ce_y || d.ce_z) throw std::invalid_argument("subpixel coefficients are not restated on the CPU");
  }

  View<T> view() {
    View<T> v;
    for (int c = 0; c < 6; ++c) v.f[c] = f[c].data();
    for (int p = 0; p < 12; ++p) v.psi[p] = psi[p].data();
    for (int a = 0; a < 3; ++a) {
      v.be[a] = coef[4 * a].data();
      v.cE[a] = coef[4 * a + 1].data();
      v.bh[a] = coef[4 * a + 2].data();
      v.cH[a] = coef[4 * a + 3].data();
      v.me[a] = me[a].data();
      v.mh[a] = mh[a].data();
      v.pe[a] = pe[a];
      v.ph[a] = ph[a];
      v.inv[a] = inv[a];
    }
    v.ce = ce.data();
    v.ch = ch;
    v.nx = nx;
    v.ny = ny;
    v.nz = nz;
    return v;
  }

  // Soft sources of one family (H: comp >= 3) at a global step; i0 shifts
  // the x index into a tile buffer of rows [i0, i0 + v.nx).
  void inject(const View<T>& v, int64_t step, bool e_side, int i0 = 0) const {
    for (const Source& s : src) {
      if ((s.comp < 3) != e_side) continue;
      const int il = s.i - i0;
      if (il < 0 || il >= v.nx) continue;
      const T val = (T)(s.amp * wave[(size_t)s.row * n_steps + step]);
      T& x = v.f[s.comp][v.at(il, s.j, s.k)];
      x = x + val;
    }
  }

  void ade_capture() {
    for (Ade& A : ade) {
      const int64_t n = (int64_t)A.cells.size();
      for (int c = 0; c < 3; ++c)
        for (int64_t q = 0; q < n; ++q) A.En[c * n + q] = f[c][A.cells[q]];
    }
  }

  void ade_apply() {
    for (Ade& A : ade) {
      const int64_t n = (int64_t)A.cells.size();
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < n; ++q) {
        for (int c = 0; c < 3; ++c) {
          T delta = 0;
          for (int p = 0; p < A.np; ++p) {
            const size_t o = ((size_t)p * 3 + c) * n + q;
            const T pc = A.P[o], pp = A.Pp[o];
            const T pn = A.a[p] * pc + A.b[p] * pp + A.c[p] * A.En[c * n + q];
            const T inc = pn - pc;
            delta = p == 0 ? inc : delta + inc;
            A.Pp[o] = pc;
            A.P[o] = pn;
          }
          if (A.np > 0) {
            T& e = f[c][A.cells[q]];
            e = e - A.inv_eps * delta;
          }
        }
      }
    }
  }

  void dft_sample(int64_t step) {
    for (Dft& D : dft) {
      if (step % D.interval != 0) continue;
      const int64_t s = step / D.interval;
      if (s >= D.n_samples) throw std::invalid_argument("DFT sample beyond the phase table");
      const pf_box& b = D.box;
      for (int q = 0; q < D.n_comp; ++q) {
        const int comp = D.comps[q];
        const double* ph = (comp >= 3 ? D.ph_h : D.ph_e).data() + 2 * (size_t)s * D.n_freq;
        const T* F = f[comp].data();
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < D.nbox; ++r) {
          const int64_t bk = r % b.count[2], bj = r / b.count[2] % b.count[1],
                        bi = r / (b.count[2] * b.count[1]);
          const double x = (double)F[(((size_t)(b.start[0] + bi * b.stride[0]) * ny) +
                                       (b.start[1] + bj * b.stride[1])) * nz +
                                      (b.start[2] + bk * b.stride[2])];
          for (int fq = 0; fq < D.n_freq; ++fq) {
            const std::complex<double> w(ph[2 * fq], ph[2 * fq + 1]);
            D.acc[((size_t)q * D.n_freq + fq) * D.nbox + r] += (w * x) * D.dt;
          }
        }
      }
    }
  }

  void step_plain(int64_t step) {
    View<T> v = view();
    for_rows(nx, [&](int a, int b) { v.h_rows(a, b); });
    inject(v, step, false);
    if (!ade.empty()) ade_capture();
    for_rows(nx, [&](int a, int b) { v.e_rows(a, b); });
    ade_apply();
    inject(v, step, true);
    dft_sample(step);
  }

  // ---- ghost-zone temporal blocking (no ADE, no DFT) --------------------
  // The 18 evolving arrays as x rows: fields (ny nz), psi with its own row
  // length; x-compacted psi rows exist only where the global map says so.
  struct Arr {
    T* base;
    size_t row;
    const int* xmap;   // null: one row per x
  };
  std::vector<Arr> arrays() {
    std::vector<Arr> out;
    for (int c = 0; c < 6; ++c) out.push_back({f[c].data(), (size_t)ny * nz, nullptr});
    const size_t rows[12] = {(size_t)pe[1] * nz, (size_t)ny * pe[2], (size_t)ny * pe[2],
                             (size_t)ny * nz,    (size_t)ny * nz,    (size_t)pe[1] * nz,
                             (size_t)ph[1] * nz, (size_t)ny * ph[2], (size_t)ny * ph[2],
                             (size_t)ny * nz,    (size_t)ny * nz,    (size_t)ph[1] * nz};
    for (int p = 0; p < 12; ++p) {
      const bool xc = p == EYX || p == EZX || p == HYX || p == HZX;
      out.push_back({psi[p].data(), rows[p], xc ? (p < 6 ? me[0].data() : mh[0].data()) : nullptr});
    }
    return out;
  }
  // Global row r of array A, or null when that x-compacted row does not exist.
  static T* row_ptr(const Arr& A, int r) {
    if (!A.xmap) return A.base + (size_t)r * A.row;
    return A.xmap[r] < 0 ? nullptr : A.base + (size_t)A.xmap[r] * A.row;
  }

  void advance_blocked(int64_t step0, int64_t n) {
    int64_t done = 0;
    while (done < n) {
      const int t = (int)std::min<int64_t>(t_block, n - done);
      const int halo = t + 2;
      const int rows = std::max(tile_rows, 2 * halo);
      if (t < 2 || nx < 2 * rows) {
        for (int64_t s = step0 + done; s < step0 + n; ++s) step_plain(s);
        return;
      }
      chunk(step0 + done, t, halo, rows);
      done += t;
    }
  }

  void chunk(int64_t step0, int t, int halo, int rows) {
    std::vector<Arr> A = arrays();
    struct Tile { int a, b, lo, hi; };
    std::vector<Tile> tiles;
    for (int a = 0; a < nx; a += rows) {
      const int b = std::min(nx, a + rows);
      tiles.push_back({a, b, std::max(0, a - halo), std::min(nx, b + halo)});
    }
    // snapshots of every array over [a - halo, a + halo) at each internal
    // tile boundary, taken before any tile writes back: snap[t][array][row]
    std::vector<std::vector<std::vector<T>>> snap(tiles.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 1; q < (int)tiles.size(); ++q) {
      const int lo = std::max(0, tiles[q].a - halo), hi = std::min(nx, tiles[q].a + halo);
      snap[q].resize(A.size());
      for (size_t ai = 0; ai < A.size(); ++ai) {
        snap[q][ai].assign((size_t)(hi - lo) * A[ai].row, T(0));
        for (int r = lo; r < hi; ++r)
          if (const T* p = row_ptr(A[ai], r))
            std::memcpy(snap[q][ai].data() + (size_t)(r - lo) * A[ai].row, p, A[ai].row * sizeof(T));
      }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < (int)tiles.size(); ++q) {
      const Tile& tl = tiles[q];
      const int ln = tl.hi - tl.lo;
      // private buffers: every array dense over the tile's rows
      std::vector<std::vector<T>> buf(A.size());
      for (size_t ai = 0; ai < A.size(); ++ai) {
        buf[ai].assign((size_t)ln * A[ai].row, T(0));
        for (int r = tl.lo; r < tl.hi; ++r) {
          T* dst = buf[ai].data() + (size_t)(r - tl.lo) * A[ai].row;
          const T* p = nullptr;
          if (r >= tl.a && r < tl.b) {
            p = row_ptr(A[ai], r);
          } else {
            const int bq = r < tl.a ? q : q + 1;   // boundary strip holding row r
            const int slo = std::max(0, tiles[bq].a - halo);
            if (A[ai].xmap == nullptr || A[ai].xmap[r] >= 0)
              p = snap[bq][ai].data() + (size_t)(r - slo) * A[ai].row;
          }
          if (p) std::memcpy(dst, p, A[ai].row * sizeof(T));
        }
      }
      // tile-local x maps: the dense buffer row index of each psi row
      std::vector<int> lme(ln), lmh(ln);
      for (int r = 0; r < ln; ++r) {
        lme[r] = me[0][tl.lo + r] >= 0 ? r : -1;
        lmh[r] = mh[0][tl.lo + r] >= 0 ? r : -1;
      }
      View<T> v = view();
      for (int c = 0; c < 6; ++c) v.f[c] = buf[c].data();
      for (int p = 0; p < 12; ++p) v.psi[p] = buf[6 + p].data();
      // x coefficients of the tile's rows
      v.be[0] = coef[0].data() + tl.lo;
      v.cE[0] = coef[1].data() + tl.lo;
      v.bh[0] = coef[2].data() + tl.lo;
      v.cH[0] = coef[3].data() + tl.lo;
      v.me[0] = lme.data();
      v.mh[0] = lmh.data();
      v.pe[0] = v.ph[0] = ln;
      v.ce = ce.data() + (size_t)tl.lo * ny * nz;
      v.nx = ln;
      for (int s = 0; s < t; ++s) {
        v.h_rows(0, ln);
        inject(v, step0 + s, false, tl.lo);
        v.e_rows(0, ln);
        inject(v, step0 + s, true, tl.lo);
      }
      // write back the core rows
      for (size_t ai = 0; ai < A.size(); ++ai)
        for (int r = tl.a; r < tl.b; ++r)
          if (T* p = row_ptr(A[ai], r))
            std::memcpy(p, buf[ai].data() + (size_t)(r - tl.lo) * A[ai].row, A[ai].row * sizeof(T));
    }
  }

  void advance(int64_t step0, int64_t n) {
    if (step0 + n > n_steps && !src.empty()) throw std::invalid_argument("steps beyond the waveform table");
    if (t_block >= 2 && ade.empty() && dft.empty()) {
      advance_blocked(step0, n);
      return;
    }
    for (int64_t s = step0; s < step0 + n; ++s) step_plain(s);
  }
};

struct Handle {
  int dtype;
  std::unique_ptr<Stepper<float>> f32;
  std::unique_ptr<Stepper<double>> f64;
};

template <typename F>
int guarded(F&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

}  // namespace

extern "C" {

const char* pfcpu_last_error(void) { return g_err.c_str(); }

// threads: OpenMP threads (0: the runtime's default); t_block >= 2 enables
// temporal blocking with tile_rows core rows per tile (when the scene has
// no dispersive media and no DFT monitors).
int pfcpu_create(const pf_desc* d, int32_t threads, int32_t t_block, int32_t tile_rows, void** out) {
  return guarded([&] {
    if (!d || !out) throw std::invalid_argument("null argument");
    if (threads > 0) omp_set_num_threads(threads);
    auto* h = new Handle{d->dtype, nullptr, nullptr};
    if (d->dtype == PF_F32) {
      h->f32.reset(new Stepper<float>(*d));
      h->f32->t_block = t_block;
      h->f32->tile_rows = tile_rows;
    } else {
      h->f64.reset(new Stepper<double>(*d));
      h->f64->t_block = t_block;
      h->f64->tile_rows = tile_rows;
    }
    *out = h;
  });
}

int pfcpu_threads(void) { return omp_get_max_threads(); }

int pfcpu_advance(void* hv, int64_t step0, int64_t n) {
  return guarded([&] {
    Handle* h = static_cast<Handle*>(hv);
    if (h->f32) h->f32->advance(step0, n);
    else h->f64->advance(step0, n);
  });
}

int pfcpu_read_field(void* hv, int32_t comp, void* dst) {
  return guarded([&] {
    Handle* h = static_cast<Handle*>(hv);
    if (comp < 0 || comp > 5) throw std::invalid_argument("component out of range");
    if (h->f32) std::memcpy(dst, h->f32->f[comp].data(), h->f32->f[comp].size() * sizeof(float));
    else std::memcpy(dst, h->f64->f[comp].data(), h->f64->f[comp].size() * sizeof(double));
  });
}

int pfcpu_dft_read(void* hv, int32_t mon, int32_t slot, double* dst) {
  return guarded([&] {
    Handle* h = static_cast<Handle*>(hv);
    const auto copy = [&](const auto& dft) {
      if (mon < 0 || mon >= (int)dft.size() || slot < 0 || slot >= dft[mon].n_comp)
        throw std::invalid_argument("DFT monitor / slot out of range");
      const auto& D = dft[mon];
      const size_t n = (size_t)D.n_freq * D.nbox;
      std::memcpy(dst, D.acc.data() + slot * n, n * 2 * sizeof(double));
    };
    if (h->f32) copy(h->f32->dft);
    else copy(h->f64->dft);
  });
}

int pfcpu_destroy(void* hv) {
  delete static_cast<Handle*>(hv);
  return 0;
}

}  // extern "C"
