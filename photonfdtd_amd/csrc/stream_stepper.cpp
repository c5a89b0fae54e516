// Beyond-HBM stepping: the domain stays in pinned host memory and x-slab
// tiles visit the device (the reference's StreamingCudaStepper,
// rust/src/cuda.rs:576-1116, chosen by rustbackend.py:264-290 when a run does
// not fit the device).
//
// A tile owns planes [a, b) and is uploaded with a halo of T planes on each
// side, [lo, hi) = [a - T, b + T) clipped to the grid; it advances T steps
// with the generic kernels (soft sources and DFT samples of its own planes
// included) and writes [a, b) back. A buffer edge spoils one plane per step
// (the H update of its last plane lacks E(i+1), the E update of its first
// lacks H(i-1)), so after T steps the owned planes hold exactly what the
// resident stepper computes: the same operations in the same order, the same
// bits. Tiles go in x order; a tile's left halo lies in the previous tile's
// owned planes, which are written back before the tile is uploaded, so those
// planes come from a snapshot taken before the chunk of T steps.
//
// Two tile buffers and three streams (uploads, kernels, downloads; default
// priority) let the next tile's upload and the previous tile's download run
// during the current tile's steps; per T steps the domain crosses the host
// link about twice, so T sets the ratio of transfer to compute.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "pfhip_host.h"
#include "pfhip_internal.h"

namespace {

using pfhip::BoxDev;
using pfhip::SourceGroupsDev;
using pfhip::StepArgs;
using pfhost::ArgError;
using pfhost::DeviceArena;
using pfhost::HipError;
using pfhost::copy_sync;

// Pinned host allocation (hipHostMalloc), freed with the stepper.
template <typename T>
struct Pinned {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    n = count;
    if (count == 0) return;
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault));
    std::memset(p, 0, count * sizeof(T));
  }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

// psi slot s (ExY ExZ EyZ EyX EzX EzY, then the H side): its CPML axis.
constexpr int kSlotAxis[12] = {1, 2, 2, 0, 0, 1, 1, 2, 2, 0, 0, 1};

template <typename real>
class Streamer final : public pf_handle {
 public:
  explicit Streamer(const pf_desc& d);
  ~Streamer() override;
  void advance(int64_t step0, int64_t n) override;
  void read_box(int comp, const pf_box& box, void* dst) override;
  void read_box_q(int comp, const pf_box& box, int bits, void* dst, double* scale) override;
  void write_field(int comp, const void* src) override;
  void dft_read(int mon, int slot, double* dst) override;
  void flux_read(double*) override {}
  void timing_enable(int) override {}
  void checkpoint_enable(int) override {
    throw ArgError{"checkpoints need the device-resident stepper (PF_FLAG_STREAM is set)"};
  }
  void adjoint(int, const pf_cotangent*, double*) override {
    throw ArgError{"the adjoint needs the device-resident stepper (PF_FLAG_STREAM is set)"};
  }
  void timing_read(double* ms, int64_t* launches) override {
    for (int c = 0; c < PF_NKCLASS; ++c) {
      if (ms) ms[c] = 0.0;
      if (launches) launches[c] = 0;
    }
  }
  void dft_flush() override;
  void info(pf_info* out) const override;
  void field_view(int, pf_view*) override {
    throw HipError{PF_E_STATE, "the streaming stepper keeps its fields in host memory; use "
                               "pf_read_field"};
  }
  void dft_view(int mon, pf_view* out) override;

 private:
  struct Tile {
    int a, b, lo, hi;
    int xrow_e[2], xrow_h[2];     // compact x psi rows of [lo, hi) and of [lo, a)
    int* map_e = nullptr;         // local compact x maps (device)
    int* map_h = nullptr;
    int pe_x = 0, ph_x = 0;       // compact x psi rows in [lo, hi)
    SourceGroupsDev src[2];       // E side, H side; cells local to the tile
    std::vector<int64_t> dft_bi;  // per monitor: first and end box x index owned
  };
  struct Buf {
    real* f[6];
    real* ce;
    real* psi[12];
    real** f_dev;
  };
  struct Dft {
    int n_comp;
    int32_t* comps_dev;
    int32_t* is_h_dev;
    BoxDev box;
    int64_t nbox;
    int n_freq, interval;
    int64_t n_samples;
    double* phase_e;
    double* phase_h;
    double dt;
    double* acc;
    real* ring;
    int K, count;
    int64_t sample0;
  };

  size_t plane_len(int s) const;      // elements of psi slot s per x plane (y/z slabs)
  size_t xrow_len() const { return (size_t)ny_ * nz_; }
  void host_rows(int s, int p0, int p1, size_t* off, size_t* len) const;
  void run_chunk(int64_t s0, int T);
  void fold(Dft& m);
  const real* host_field(int c) const { return hf_.p + (size_t)c * ncell_; }

  int nx_, ny_, nz_;
  int64_t ncell_;
  bool exact_;
  int64_t n_steps_;
  int T_, rows_;
  hipStream_t up_ = nullptr, run_ = nullptr, down_ = nullptr;
  hipEvent_t ev_up_[2] = {}, ev_run_[2] = {}, ev_down_[2] = {};
  DeviceArena arena_;
  // host state
  Pinned<real> hf_;          // six components back to back
  Pinned<real> hce_;
  Pinned<real> hpsi_;        // twelve slots back to back
  size_t hpsi_off_[12] = {};
  Pinned<real> strip_;       // left-halo snapshot of every tile for one chunk
  std::vector<size_t> strip_off_;
  std::vector<int> me_x_, mh_x_;   // global compact x maps
  StepArgs<real> g_;         // global y/z tables and scalars
  std::vector<Tile> tiles_;
  Buf buf_[2];
  std::vector<Dft> dft_;
  double* qpart_ = nullptr;
  void* qbuf_ = nullptr;
  size_t qbuf_bytes_ = 0;
  real* qsrc_ = nullptr;
  size_t qsrc_len_ = 0;
};

template <typename real>
size_t Streamer<real>::plane_len(int s) const {
  const int ax = kSlotAxis[s];
  const int pc = s < 6 ? g_.pe[ax] : g_.ph[ax];
  return ax == 1 ? (size_t)pc * nz_ : (size_t)ny_ * pc;
}

// Host range of psi slot s covering planes [p0, p1).
template <typename real>
void Streamer<real>::host_rows(int s, int p0, int p1, size_t* off, size_t* len) const {
  if (kSlotAxis[s] == 0) {
    const std::vector<int>& m = s < 6 ? me_x_ : mh_x_;
    int r0 = 0, r1 = 0;
    for (int i = 0; i < p1; ++i)
      if (m[i] >= 0) {
        if (i < p0) ++r0;
        ++r1;
      }
    *off = hpsi_off_[s] + (size_t)r0 * xrow_len();
    *len = (size_t)(r1 - r0) * xrow_len();
  } else {
    *off = hpsi_off_[s] + (size_t)p0 * plane_len(s);
    *len = (size_t)(p1 - p0) * plane_len(s);
  }
}

template <typename real>
Streamer<real>::Streamer(const pf_desc& d) {
  for (int a = 0; a < 3; ++a) {
    shape[a] = d.shape[a];
    if (d.shape[a] < 1) throw ArgError{"shape must be positive"};
  }
  nx_ = (int)d.shape[0];
  ny_ = (int)d.shape[1];
  nz_ = (int)d.shape[2];
  ncell_ = (int64_t)nx_ * ny_ * nz_;
  if (nx_ < 2) throw ArgError{"streaming needs at least two x planes"};
  if ((int64_t)ny_ * nz_ >= (int64_t(1) << 31) || ncell_ >= (int64_t(1) << 40))
    throw ArgError{"grid extent beyond the streaming stepper's limits"};
  if (!d.ce) throw ArgError{"ce is required"};
  for (int a = 0; a < 3; ++a)
    if (!d.b_e[a] || !d.c_e[a] || !d.b_h[a] || !d.c_h[a])
      throw ArgError{"CPML coefficient arrays are required on every axis"};
  if (d.n_ade > 0) throw ArgError{"the streaming stepper does not step dispersive media"};
  if (d.n_flux > 0) throw ArgError{"the streaming stepper does not accumulate flux monitors"};
  if (d.ce_y || d.ce_z) throw ArgError{"the streaming stepper does not take subpixel coefficients"};
  if (d.n_steps < 0 || d.n_src < 0 || d.n_wave < 0 || d.n_dft < 0)
    throw ArgError{"negative table length"};
  if (d.n_src > 0 && (!d.src_comp || !d.src_cell || !d.src_amp || !d.src_wave || !d.wave))
    throw ArgError{"source tables missing"};
  exact_ = d.exact != 0;
  n_steps_ = d.n_steps;
  T_ = d.t_block > 0 ? d.t_block : 32;
  if (T_ > 1024) throw ArgError{"t_block must be at most 1024"};

  HIP_CHECK(hipStreamCreateWithFlags(&up_, hipStreamNonBlocking));
  HIP_CHECK(hipStreamCreateWithFlags(&run_, hipStreamNonBlocking));
  HIP_CHECK(hipStreamCreateWithFlags(&down_, hipStreamNonBlocking));
  for (int q = 0; q < 2; ++q) {
    HIP_CHECK(hipEventCreateWithFlags(&ev_up_[q], hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&ev_run_[q], hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&ev_down_[q], hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(ev_down_[q], down_));
  }

  // global tables: CPML coefficients (x rows are offset per tile) and maps
  std::memset(&g_, 0, sizeof(g_));
  g_.ny = ny_;
  g_.nz = nz_;
  g_.nyz = (int64_t)ny_ * nz_;
  for (int a = 0; a < 3; ++a) {
    g_.d[a] = (real)d.cell[a];
    g_.inv_d[a] = (real)(1.0 / d.cell[a]);
  }
  g_.ch = (real)d.ch;
  const int n_axis[3] = {nx_, ny_, nz_};
  for (int a = 0; a < 3; ++a) {
    const int n = n_axis[a];
    std::vector<real> tbl(4 * (size_t)n);
    std::vector<int> me(n, -1), mh(n, -1);
    int pe = 0, ph = 0;
    for (int i = 0; i < n; ++i) {
      tbl[i] = (real)d.b_e[a][i];
      tbl[n + i] = (real)d.c_e[a][i];
      tbl[2 * n + i] = (real)d.b_h[a][i];
      tbl[3 * n + i] = (real)d.c_h[a][i];
      if (d.c_e[a][i] != 0.0) me[i] = pe++;
      if (d.c_h[a][i] != 0.0) mh[i] = ph++;
    }
    real* t = arena_.upload(tbl.data(), tbl.size(), run_);
    g_.b_e[a] = t;
    g_.c_e[a] = t + n;
    g_.b_h[a] = t + 2 * n;
    g_.c_h[a] = t + 3 * n;
    g_.pe[a] = pe;
    g_.ph[a] = ph;
    if (a == 0) {
      me_x_ = me;
      mh_x_ = mh;
    } else {
      g_.map_e[a] = arena_.upload(me.data(), n, run_);
      g_.map_h[a] = arena_.upload(mh.data(), n, run_);
    }
  }

  // host state
  hf_.alloc(6 * (size_t)ncell_);
  hce_.alloc((size_t)ncell_);
  std::memcpy(hce_.p, d.ce, (size_t)ncell_ * sizeof(real));
  size_t psi_total = 0;
  for (int s = 0; s < 12; ++s) {
    hpsi_off_[s] = psi_total;
    const int ax = kSlotAxis[s];
    psi_total += ax == 0 ? (size_t)(s < 6 ? g_.pe[0] : g_.ph[0]) * xrow_len()
                         : (size_t)nx_ * plane_len(s);
  }
  hpsi_.alloc(psi_total);

  // tiles: owned planes [a, b), halo T on each side
  size_t plane_bytes = (size_t)ny_ * nz_ * sizeof(real);
  size_t per_plane = 7 * (size_t)ny_ * nz_;   // six fields and ce
  for (int s = 0; s < 12; ++s) per_plane += kSlotAxis[s] == 0 ? xrow_len() : plane_len(s);
  if (d.tile_rows > 0) {
    rows_ = d.tile_rows;
  } else {
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    const size_t fit = free_b / 4 / (2 * per_plane * sizeof(real));   // two buffers in a quarter
    rows_ = (int)std::max<int64_t>(2, (int64_t)fit - 2 * T_);
  }
  (void)plane_bytes;
  rows_ = std::min(rows_, nx_);
  for (int a = 0; a < nx_; a += rows_) {
    Tile t;
    t.a = a;
    t.b = std::min(nx_, a + rows_);
    t.lo = std::max(0, a - T_);
    t.hi = std::min(nx_, t.b + T_);
    tiles_.push_back(t);
  }
  int rmax = 0, xe_max = 0, xh_max = 0;
  for (Tile& t : tiles_) {
    rmax = std::max(rmax, t.hi - t.lo);
    std::vector<int> le(t.hi - t.lo, -1), lh(t.hi - t.lo, -1);
    int ne = 0, nh = 0, ne_a = 0, nh_a = 0, e0 = 0, h0 = 0;
    for (int i = 0; i < t.lo; ++i) {
      e0 += me_x_[i] >= 0;
      h0 += mh_x_[i] >= 0;
    }
    for (int i = t.lo; i < t.hi; ++i) {
      if (me_x_[i] >= 0) {
        le[i - t.lo] = ne++;
        if (i < t.a) ++ne_a;
      }
      if (mh_x_[i] >= 0) {
        lh[i - t.lo] = nh++;
        if (i < t.a) ++nh_a;
      }
    }
    t.pe_x = ne;
    t.ph_x = nh;
    t.xrow_e[0] = e0;
    t.xrow_e[1] = ne_a;
    t.xrow_h[0] = h0;
    t.xrow_h[1] = nh_a;
    xe_max = std::max(xe_max, ne);
    xh_max = std::max(xh_max, nh);
    t.map_e = arena_.upload(le.data(), le.size(), run_);
    t.map_h = arena_.upload(lh.data(), lh.size(), run_);
  }
  for (int q = 0; q < 2; ++q) {
    Buf& B = buf_[q];
    real* lvl = arena_.zeros<real>(6 * (size_t)rmax * ny_ * nz_, run_);
    for (int c = 0; c < 6; ++c) B.f[c] = lvl + (size_t)c * rmax * ny_ * nz_;
    B.ce = arena_.zeros<real>((size_t)rmax * ny_ * nz_, run_);
    for (int s = 0; s < 12; ++s) {
      const size_t len = kSlotAxis[s] == 0 ? (size_t)(s < 6 ? xe_max : xh_max) * xrow_len()
                                           : (size_t)rmax * plane_len(s);
      B.psi[s] = arena_.zeros<real>(std::max<size_t>(len, 1), run_);
    }
    B.f_dev = arena_.upload(B.f, 6, run_);
  }
  // left-halo snapshots (pinned so the uploads can run asynchronously)
  size_t strip_total = 0;
  for (const Tile& t : tiles_) {
    strip_off_.push_back(strip_total);
    size_t len = 6 * (size_t)(t.a - t.lo) * ny_ * nz_;
    for (int s = 0; s < 12; ++s) {
      size_t off, l;
      host_rows(s, t.lo, t.a, &off, &l);
      len += l;
    }
    strip_total += len;
  }
  strip_.alloc(strip_total);

  // sources: the groups of the resident stepper (consecutive equal (comp,
  // cell) entries), each tile holding those of its planes, cells local
  std::vector<int32_t> gcomp[2], gbeg[2];
  std::vector<int64_t> gcell[2];
  std::vector<double> amp[2];
  std::vector<int32_t> wrow[2];
  for (int s = 0; s < d.n_src; ++s) {
    const int comp = d.src_comp[s];
    if (comp < 0 || comp > 5) throw ArgError{"source component out of range"};
    if (d.src_cell[s] < 0 || d.src_cell[s] >= ncell_) throw ArgError{"source cell out of range"};
    if (d.src_wave[s] < 0 || d.src_wave[s] >= d.n_wave) throw ArgError{"source waveform row out of range"};
    const int side = comp >= 3 ? 1 : 0;
    if (gcomp[side].empty() || gcomp[side].back() != comp || gcell[side].back() != d.src_cell[s]) {
      gcomp[side].push_back(comp);
      gcell[side].push_back(d.src_cell[s]);
      gbeg[side].push_back((int32_t)amp[side].size());
    }
    amp[side].push_back(d.src_amp[s]);
    wrow[side].push_back(d.src_wave[s]);
  }
  double* wave = arena_.upload(d.wave, (size_t)d.n_wave * d.n_steps, run_);
  const int64_t nyz = (int64_t)ny_ * nz_;
  for (Tile& t : tiles_)
    for (int side = 0; side < 2; ++side) {
      std::vector<int32_t> c, b;
      std::vector<int64_t> cell;
      std::vector<double> am;
      std::vector<int32_t> wr;
      for (size_t g = 0; g < gcomp[side].size(); ++g) {
        const int64_t x = gcell[side][g] / nyz;
        if (x < t.lo || x >= t.hi) continue;
        c.push_back(gcomp[side][g]);
        cell.push_back(gcell[side][g] - (int64_t)t.lo * nyz);
        b.push_back((int32_t)am.size());
        const int32_t e = g + 1 < gbeg[side].size() ? gbeg[side][g + 1] : (int32_t)amp[side].size();
        for (int32_t k = gbeg[side][g]; k < e; ++k) {
          am.push_back(amp[side][k]);
          wr.push_back(wrow[side][k]);
        }
      }
      SourceGroupsDev& G = t.src[side];
      G.n_groups = (int32_t)c.size();
      if (G.n_groups == 0) continue;
      b.push_back((int32_t)am.size());
      G.comp = arena_.upload(c.data(), c.size(), run_);
      G.cell = arena_.upload(cell.data(), cell.size(), run_);
      G.begin = arena_.upload(b.data(), b.size(), run_);
      G.amp = arena_.upload(am.data(), am.size(), run_);
      G.wave_row = arena_.upload(wr.data(), wr.size(), run_);
      G.wave = wave;
    }

  // DFT monitors: rings and accumulators as in the resident stepper; each
  // tile gathers the box x indices of its owned planes
  for (int m = 0; m < d.n_dft; ++m) {
    const pf_dft_desc& s = d.dft[m];
    if (s.n_comp < 1 || s.n_comp > 6 || s.n_freq < 1 || s.interval < 1 || s.n_samples < 0)
      throw ArgError{"bad DFT monitor description"};
    Dft st;
    st.n_comp = s.n_comp;
    int32_t comps[6], is_h[6];
    for (int c = 0; c < s.n_comp; ++c) {
      if (s.comps[c] < 0 || s.comps[c] > 5) throw ArgError{"DFT component out of range"};
      comps[c] = s.comps[c];
      is_h[c] = s.comps[c] >= 3;
    }
    st.comps_dev = arena_.upload(comps, s.n_comp, run_);
    st.is_h_dev = arena_.upload(is_h, s.n_comp, run_);
    st.box = pfhost::check_box(s.box, shape);
    st.nbox = st.box.count[0] * st.box.count[1] * st.box.count[2];
    st.n_freq = s.n_freq;
    st.interval = s.interval;
    st.n_samples = s.n_samples;
    st.phase_e = arena_.upload(s.phase_e, (size_t)s.n_samples * s.n_freq * 2, run_);
    st.phase_h = arena_.upload(s.phase_h, (size_t)s.n_samples * s.n_freq * 2, run_);
    st.dt = s.dt;
    st.acc = arena_.zeros<double>((size_t)s.n_comp * s.n_freq * st.nbox * 2, run_);
    const int64_t sample_bytes = (int64_t)s.n_comp * st.nbox * (int64_t)sizeof(real);
    const int64_t budget = 1ll << 30;
    st.K = (int)std::max<int64_t>(
        1, std::min<int64_t>(std::max(128, T_), budget / std::max<int64_t>(sample_bytes, 1)));
    st.ring = static_cast<real*>(arena_.alloc((size_t)st.K * sample_bytes));
    st.count = 0;
    st.sample0 = 0;
    dft_.push_back(st);
    for (Tile& t : tiles_) {
      // owned box x indices: start + bi * stride in [a, b)
      const int64_t x0 = st.box.start[0], sx = st.box.stride[0], cx = st.box.count[0];
      int64_t b0 = t.a <= x0 ? 0 : (t.a - x0 + sx - 1) / sx;
      int64_t b1 = t.b <= x0 ? 0 : (t.b - x0 + sx - 1) / sx;
      b0 = std::min(b0, cx);
      b1 = std::min(b1, cx);
      t.dft_bi.push_back(b0);
      t.dft_bi.push_back(b1);
    }
  }
  HIP_CHECK(hipStreamSynchronize(run_));
}

template <typename real>
Streamer<real>::~Streamer() {
  for (hipStream_t s : {up_, run_, down_})
    if (s) (void)hipStreamSynchronize(s);
  for (int q = 0; q < 2; ++q)
    for (hipEvent_t e : {ev_up_[q], ev_run_[q], ev_down_[q]})
      if (e) (void)hipEventDestroy(e);
  arena_.release();
  for (hipStream_t s : {up_, run_, down_})
    if (s) (void)hipStreamDestroy(s);
}

template <typename real>
void Streamer<real>::advance(int64_t step0, int64_t n) {
  if (step0 < 0 || n < 0) throw ArgError{"negative step range"};
  bool any_src = false;
  for (const Tile& t : tiles_) any_src = any_src || t.src[0].n_groups || t.src[1].n_groups;
  if (any_src && step0 + n > n_steps_) throw ArgError{"step range exceeds the source waveform table"};
  for (int64_t s = step0; s < step0 + n;) {
    int T = (int)std::min<int64_t>(T_, step0 + n - s);
    // every sample of the chunk must fit its monitor's ring
    for (const Dft& m : dft_) {
      auto samples = [&](int len) {
        int c = 0;
        for (int64_t t = s; t < s + len; ++t) c += t % m.interval == 0;
        return c;
      };
      while (T > 1 && samples(T) > m.K) --T;
    }
    for (Dft& m : dft_) {
      int c = 0;
      for (int64_t t = s; t < s + T; ++t) c += t % m.interval == 0;
      if (c > 0) {
        const int64_t first = (s + m.interval - 1) / m.interval;
        if (first + c > m.n_samples) throw ArgError{"DFT sample beyond the phase table"};
        if (m.count > 0 && first != m.sample0 + m.count) throw ArgError{"DFT samples must be consecutive"};
        if (m.count + c > m.K) fold(m);
      }
    }
    run_chunk(s, T);
    s += T;
  }
  HIP_CHECK(hipStreamSynchronize(run_));
}

template <typename real>
void Streamer<real>::run_chunk(int64_t s0, int T) {
  const size_t nyz = (size_t)ny_ * nz_;
  // snapshot every tile's left halo before any tile writes back
  for (size_t k = 0; k < tiles_.size(); ++k) {
    const Tile& t = tiles_[k];
    real* dst = strip_.p + strip_off_[k];
    const size_t fl = (size_t)(t.a - t.lo) * nyz;
    for (int c = 0; c < 6; ++c, dst += fl)
      std::memcpy(dst, hf_.p + (size_t)c * ncell_ + (size_t)t.lo * nyz, fl * sizeof(real));
    for (int s = 0; s < 12; ++s) {
      size_t off, len;
      host_rows(s, t.lo, t.a, &off, &len);
      std::memcpy(dst, hpsi_.p + off, len * sizeof(real));
      dst += len;
    }
  }
  // samples of each monitor in the chunk and their first ring slot
  std::vector<int> first_slot(dft_.size()), n_samp(dft_.size());
  for (size_t m = 0; m < dft_.size(); ++m) {
    Dft& M = dft_[m];
    int c = 0;
    for (int64_t t = s0; t < s0 + T; ++t) c += t % M.interval == 0;
    n_samp[m] = c;
    first_slot[m] = M.count;
    if (c > 0 && M.count == 0) M.sample0 = (s0 + M.interval - 1) / M.interval;
  }

  for (size_t k = 0; k < tiles_.size(); ++k) {
    const Tile& t = tiles_[k];
    const int q = (int)(k & 1);
    Buf& B = buf_[q];
    const size_t nl = (size_t)(t.hi - t.lo);
    // ---- upload: left halo from the snapshot, the rest from the host state ----
    HIP_CHECK(hipStreamWaitEvent(up_, ev_down_[q], 0));
    const real* sp = strip_.p + strip_off_[k];
    const size_t fl = (size_t)(t.a - t.lo) * nyz;
    for (int c = 0; c < 6; ++c) {
      if (fl) HIP_CHECK(hipMemcpyAsync(B.f[c], sp + c * fl, fl * sizeof(real), hipMemcpyHostToDevice, up_));
      HIP_CHECK(hipMemcpyAsync(B.f[c] + fl, hf_.p + (size_t)c * ncell_ + (size_t)t.a * nyz,
                               (nl * nyz - fl) * sizeof(real), hipMemcpyHostToDevice, up_));
    }
    sp += 6 * fl;
    for (int s = 0; s < 12; ++s) {
      size_t off_h, len_h, off_r, len_r;
      host_rows(s, t.lo, t.a, &off_h, &len_h);   // snapshot part
      host_rows(s, t.a, t.hi, &off_r, &len_r);   // host part
      if (len_h) HIP_CHECK(hipMemcpyAsync(B.psi[s], sp, len_h * sizeof(real), hipMemcpyHostToDevice, up_));
      if (len_r)
        HIP_CHECK(hipMemcpyAsync(B.psi[s] + len_h, hpsi_.p + off_r, len_r * sizeof(real),
                                 hipMemcpyHostToDevice, up_));
      sp += len_h;
    }
    HIP_CHECK(hipMemcpyAsync(B.ce, hce_.p + (size_t)t.lo * nyz, nl * nyz * sizeof(real),
                             hipMemcpyHostToDevice, up_));
    HIP_CHECK(hipEventRecord(ev_up_[q], up_));

    // ---- T steps on the tile ----
    HIP_CHECK(hipStreamWaitEvent(run_, ev_up_[q], 0));
    StepArgs<real> a = g_;
    a.nx = t.hi - t.lo;
    for (int c = 0; c < 6; ++c) a.f[c] = B.f[c];
    a.ce = B.ce;
    for (int s = 0; s < 12; ++s) a.psi[s] = B.psi[s];
    a.b_e[0] = g_.b_e[0] + t.lo;
    a.c_e[0] = g_.c_e[0] + t.lo;
    a.b_h[0] = g_.b_h[0] + t.lo;
    a.c_h[0] = g_.c_h[0] + t.lo;
    a.map_e[0] = t.map_e;
    a.map_h[0] = t.map_h;
    a.pe[0] = t.pe_x;
    a.ph[0] = t.ph_x;
    std::vector<int> slot(first_slot);
    for (int64_t st = s0; st < s0 + T; ++st) {
      pfhip::launch_generic_h<real>(a, exact_, run_);
      pfhip::launch_inject<real>(B.f_dev, t.src[1], n_steps_, st, run_);
      pfhip::launch_generic_e<real>(a, exact_, run_);
      pfhip::launch_inject<real>(B.f_dev, t.src[0], n_steps_, st, run_);
      for (size_t m = 0; m < dft_.size(); ++m) {
        Dft& M = dft_[m];
        if (st % M.interval != 0) continue;
        const int64_t b0 = t.dft_bi[2 * m], b1 = t.dft_bi[2 * m + 1];
        if (b1 > b0) {
          BoxDev part = M.box;
          part.start[0] = M.box.start[0] + b0 * M.box.stride[0] - t.lo;
          part.count[0] = b1 - b0;
          real* out = M.ring + (int64_t)slot[m] * M.n_comp * M.nbox +
                      b0 * M.box.count[1] * M.box.count[2];
          pfhip::launch_gather_box<real>(B.f_dev, M.comps_dev, M.n_comp, part, (int64_t)nyz, nz_, 0,
                                         out, run_, M.nbox);
        }
        ++slot[m];
      }
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ev_run_[q], run_));

    // ---- download the owned planes ----
    HIP_CHECK(hipStreamWaitEvent(down_, ev_run_[q], 0));
    const size_t own0 = (size_t)(t.a - t.lo) * nyz, own = (size_t)(t.b - t.a) * nyz;
    for (int c = 0; c < 6; ++c)
      HIP_CHECK(hipMemcpyAsync(hf_.p + (size_t)c * ncell_ + (size_t)t.a * nyz, B.f[c] + own0,
                               own * sizeof(real), hipMemcpyDeviceToHost, down_));
    for (int s = 0; s < 12; ++s) {
      size_t off_h, len_h, off_o, len_o;
      host_rows(s, t.lo, t.a, &off_h, &len_h);
      host_rows(s, t.a, t.b, &off_o, &len_o);
      if (len_o)
        HIP_CHECK(hipMemcpyAsync(hpsi_.p + off_o, B.psi[s] + len_h, len_o * sizeof(real),
                                 hipMemcpyDeviceToHost, down_));
    }
    HIP_CHECK(hipEventRecord(ev_down_[q], down_));
  }
  for (size_t m = 0; m < dft_.size(); ++m) dft_[m].count += n_samp[m];
  HIP_CHECK(hipStreamSynchronize(down_));
  HIP_CHECK(hipStreamSynchronize(up_));
}

template <typename real>
void Streamer<real>::fold(Dft& m) {
  if (m.count == 0) return;
  pfhip::launch_dft_fold<real>(m.ring, m.nbox, m.n_comp, m.is_h_dev, m.count, m.sample0, m.phase_e,
                               m.phase_h, m.n_freq, m.dt, m.acc, run_);
  m.sample0 += m.count;
  m.count = 0;
}

template <typename real>
void Streamer<real>::dft_flush() {
  for (Dft& m : dft_) fold(m);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(run_));
}

template <typename real>
void Streamer<real>::dft_read(int mon, int slot, double* dst) {
  if (mon < 0 || mon >= (int)dft_.size()) throw ArgError{"DFT monitor index out of range"};
  Dft& m = dft_[mon];
  if (slot < 0 || slot >= m.n_comp) throw ArgError{"DFT component slot out of range"};
  if (!dst) throw ArgError{"null destination"};
  fold(m);
  const size_t n = (size_t)m.n_freq * m.nbox * 2;
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(dst, m.acc + (size_t)slot * n, n * sizeof(double), hipMemcpyDeviceToHost, run_));
  HIP_CHECK(hipStreamSynchronize(run_));
}

template <typename real>
void Streamer<real>::dft_view(int mon, pf_view* out) {
  if (mon < 0 || mon >= (int)dft_.size()) throw ArgError{"DFT monitor index out of range"};
  Dft& m = dft_[mon];
  fold(m);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(run_));
  std::memset(out, 0, sizeof(*out));
  out->ptr = m.acc;
  out->dtype = PF_F64;
  out->ndim = 6;
  const int64_t dims[6] = {m.n_comp, m.n_freq, m.box.count[0], m.box.count[1], m.box.count[2], 2};
  int64_t stride = 1;
  for (int a = 5; a >= 0; --a) {
    out->shape[a] = dims[a];
    out->strides[a] = stride;
    stride *= dims[a];
  }
  out->stream = (uint64_t)(uintptr_t)run_;
}

template <typename real>
void Streamer<real>::read_box(int comp, const pf_box& box, void* dst) {
  if (comp < 0 || comp > 5) throw ArgError{"component out of range"};
  if (!dst) throw ArgError{"null destination"};
  const BoxDev b = pfhost::check_box(box, shape);
  const real* F = host_field(comp);
  real* out = static_cast<real*>(dst);
  for (int64_t i = 0; i < b.count[0]; ++i)
    for (int64_t j = 0; j < b.count[1]; ++j) {
      const real* row = F + ((b.start[0] + i * b.stride[0]) * ny_ + b.start[1] + j * b.stride[1]) * nz_;
      for (int64_t k = 0; k < b.count[2]; ++k) *out++ = row[b.start[2] + k * b.stride[2]];
    }
}

template <typename real>
void Streamer<real>::read_box_q(int comp, const pf_box& box, int bits, void* dst, double* scale) {
  if (!dst || !scale) throw ArgError{"null destination"};
  if (bits != 8 && bits != 16) throw ArgError{"bits must be 8 or 16"};
  const BoxDev b = pfhost::check_box(box, shape);
  const int64_t nbox = b.count[0] * b.count[1] * b.count[2];
  std::vector<real> host((size_t)nbox);
  read_box(comp, box, host.data());
  // the device quantisation of the resident stepper, on the uploaded box
  constexpr int kMaxPart = 1024;
  if (!qpart_) qpart_ = static_cast<double*>(arena_.alloc((kMaxPart + 1) * sizeof(double)));
  if ((size_t)nbox > qsrc_len_) {
    qsrc_ = static_cast<real*>(arena_.alloc((size_t)nbox * sizeof(real)));
    qsrc_len_ = (size_t)nbox;
  }
  const size_t bytes = (size_t)nbox * (bits / 8);
  if (bytes > qbuf_bytes_) {
    qbuf_ = arena_.alloc(bytes);
    qbuf_bytes_ = bytes;
  }
  copy_sync(qsrc_, host.data(), (size_t)nbox * sizeof(real), hipMemcpyHostToDevice, run_);
  pfhip::launch_quantize<real>(qsrc_, nbox, bits, qpart_, kMaxPart, qbuf_, qpart_ + kMaxPart, run_);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(dst, qbuf_, bytes, hipMemcpyDeviceToHost, run_));
  HIP_CHECK(hipMemcpyAsync(scale, qpart_ + kMaxPart, sizeof(double), hipMemcpyDeviceToHost, run_));
  HIP_CHECK(hipStreamSynchronize(run_));
}

template <typename real>
void Streamer<real>::write_field(int comp, const void* src) {
  if (comp < 0 || comp > 5) throw ArgError{"component out of range"};
  if (!src) throw ArgError{"null source"};
  std::memcpy(hf_.p + (size_t)comp * ncell_, src, (size_t)ncell_ * sizeof(real));
}

template <typename real>
void Streamer<real>::info(pf_info* out) const {
  std::memset(out, 0, sizeof(*out));
  out->path = 4;
  out->zrow = nz_;
  out->tiles_bulk = (int64_t)tiles_.size();
  out->cells_bulk = ncell_;
  for (const Dft& m : dft_) out->dft_ring = std::max<int64_t>(out->dft_ring, m.K);
}

}  // namespace

namespace pfhost {

pf_handle* make_streamer(const pf_desc& d) {
  if (d.dtype == PF_F32) return new Streamer<float>(d);
  if (d.dtype == PF_F64) return new Streamer<double>(d);
  throw ArgError{"dtype must be PF_F32 or PF_F64"};
}

}  // namespace pfhost
