// libpfhip: C ABI (include/pfhip.h) over the gfx950 FDTD kernels.
//
// A handle owns every device buffer of one simulation -- fields, the E
// coefficient, compact CPML psi, source tables, ADE state, DFT accumulators
// and flux totals -- on its own HIP stream (default priority). The host
// crosses into device memory only at pf_create (upload), pf_read_* /
// pf_dft_read (download) and pf_write_field, mirroring the reference's
// VRAM-resident CudaStepper (rust/src/cuda.rs:271-574) but with ADE, DFT and
// flux accumulation kept on the device as well.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <tuple>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/pfhip.h"
#include "pfhip_internal.h"
#include "pfhip_host.h"

namespace {

thread_local std::string g_last_error = "";

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

using pfhost::ArgError;
using pfhost::DeviceArena;
using pfhost::HipError;
using pfhost::copy_sync;

// True when the nodes with map >= 0 are exactly [0, lo) U [hi, n) with the
// compact index running in node order; stores (lo, hi).
bool two_slabs(const std::vector<int>& map, int out[2]) {
  const int n = (int)map.size();
  int lo = 0;
  while (lo < n && map[lo] >= 0) ++lo;
  int hi = lo;
  while (hi < n && map[hi] < 0) ++hi;
  for (int i = hi; i < n; ++i)
    if (map[i] < 0) return false;
  for (int i = 0; i < n; ++i) {
    const bool in = i < lo || i >= hi;
    if (in != (map[i] >= 0)) return false;
    if (in && map[i] != (i < lo ? i : lo + (i - hi))) return false;
  }
  out[0] = lo;
  out[1] = hi;
  return true;
}

// Split [0, n) into segments of at most `len` nodes that never straddle the
// bulk range [b0, b1): (start, count, in_bulk).
struct Seg {
  int start, count;
  bool bulk;
};
std::vector<Seg> split_axis(int n, int b0, int b1, int len, bool widen) {
  std::vector<Seg> out;
  auto emit = [&](int lo, int hi, bool bulk) {
    for (int s = lo; s < hi; s += len) out.push_back({s, std::min(len, hi - s), bulk});
  };
  b0 = std::max(0, std::min(b0, n));
  b1 = std::max(b0, std::min(b1, n));
  if (widen) {
    // edge segments rounded up to whole tiles, so CPML tiles use every lane/row
    const int lo = b0 > 0 ? std::min(n, (b0 + len - 1) / len * len) : 0;
    const int hi = b1 < n ? std::max(lo, n - (n - b1 + len - 1) / len * len) : n;
    b0 = lo;
    b1 = std::max(lo, hi);
  }
  emit(0, b0, false);
  emit(b0, b1, true);
  emit(b1, n, false);
  return out;
}

// [0, n) as [0, b0) U [b0, b1) U [b1, n), each range cut into the fewest
// segments of at most len nodes, of nearly equal size.
std::vector<Seg> split_even(int n, int b0, int b1, int len) {
  std::vector<Seg> out;
  b0 = std::max(0, std::min(b0, n));
  b1 = std::max(b0, std::min(b1, n));
  auto emit = [&](int lo, int hi, bool bulk) {
    if (hi <= lo) return;
    const int k = (hi - lo + len - 1) / len;
    for (int i = 0; i < k; ++i) {
      const int s = lo + (int)((int64_t)(hi - lo) * i / k);
      const int e = lo + (int)((int64_t)(hi - lo) * (i + 1) / k);
      out.push_back({s, e - s, bulk});
    }
  };
  emit(0, b0, false);
  emit(b0, b1, true);
  emit(b1, n, false);
  return out;
}

// Segments of `len` nodes aligned to 0 (the last one shorter), bulk when
// inside [b0, b1). The dense-row march cuts y and z this way: a segment that
// reaches into a CPML slab keeps its full length and runs with the kernel
// variant of that axis, so the CPML tiles stay as full as the bulk ones.
std::vector<Seg> split_aligned(int n, int b0, int b1, int len) {
  std::vector<Seg> out;
  for (int s = 0; s < n; s += len) {
    const int c = std::min(len, n - s);
    out.push_back({s, c, s >= b0 && s + c <= b1});
  }
  return out;
}

// Dense rows [nx][ny][nz] -> padded rows [nx][ny][zrow], zpad nodes per
// block (pad slots zero).
template <typename real>
std::vector<real> pad_rows(const real* src, int nx, int ny, int nz, int zrow, int zpad) {
  std::vector<real> out((size_t)nx * ny * zrow, real(0));
  for (int64_t r = 0; r < (int64_t)nx * ny; ++r)
    for (int k = 0; k < nz; ++k) out[(size_t)(r * zrow + pfhip::zslot(k, zpad))] = src[r * nz + k];
  return out;
}
// Dense cell index -> index in padded rows.
int64_t pad_cell(int64_t c, int nz, int zrow, int zpad) {
  return (c / nz) * zrow + pfhip::zslot(c % nz, zpad);
}

// Order tiles so that the blocks the hardware sends to one XCD (block ids
// congruent mod 8) are spatially contiguous; pad with empty tiles.
std::vector<pfhip::MarchTile> xcd_order(const std::vector<pfhip::MarchTile>& t) {
  const size_t g = (t.size() + 7) / 8;
  std::vector<pfhip::MarchTile> out(8 * g, pfhip::MarchTile{0, 0, 0, 0, 0, 0, 0, 0});
  for (size_t idx = 0; idx < t.size(); ++idx) {
    const size_t xcd = idx / g, q = idx % g;
    out[8 * q + xcd] = t[idx];
  }
  return out;
}

// Alternative order: every x-chunk is split into 8 contiguous bands (one per
// XCD), and each XCD takes its band of chunk 0, then of chunk 1, ..., so all
// XCDs sweep the same x-chunk at the same time.
std::vector<pfhip::MarchTile> xcd_band_order(const std::vector<pfhip::MarchTile>& t) {
  std::vector<std::vector<pfhip::MarchTile>> queue(8);
  size_t a = 0;
  while (a < t.size()) {
    size_t b = a;
    while (b < t.size() && t[b].i0 == t[a].i0) ++b;   // one x-chunk
    const size_t m = b - a;
    for (int x = 0; x < 8; ++x)
      for (size_t q = a + m * x / 8; q < a + m * (x + 1) / 8; ++q) queue[x].push_back(t[q]);
    a = b;
  }
  size_t g = 0;
  for (auto& q : queue) g = std::max(g, q.size());
  std::vector<pfhip::MarchTile> out(8 * g, pfhip::MarchTile{0, 0, 0, 0, 0, 0, 0, 0});
  for (int x = 0; x < 8; ++x)
    for (size_t q = 0; q < queue[x].size(); ++q) out[8 * q + x] = queue[x][q];
  return out;
}

using pfhip::BoxDev;
using pfhip::SourceGroupsDev;
using pfhip::StepArgs;

using pfhost::check_box;

}  // namespace

namespace {

template <typename real>
class Stepper final : public pf_handle {
 public:
  // zpad_default: the march's row layout when PHOTONFDTD_HIP_ZPAD does not
  // name one (0: chosen by grid size)
  explicit Stepper(const pf_desc& d, int zpad_default = 0);
  ~Stepper() override {
    if (stream_) {
      (void)hipStreamSynchronize(stream_);
      for (Timed& t : pending_) {
        if (!t.shared_start) (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
      }
      for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
      for (hipEvent_t e : fork_ev_)
        if (e) (void)hipEventDestroy(e);
      if (side_) {
        (void)hipStreamSynchronize(side_);
        (void)hipStreamDestroy(side_);
      }
      arena_.release();
      if (own_stream_) (void)hipStreamDestroy(stream_);
    }
  }
  void advance(int64_t step0, int64_t n) override;
  void read_box(int comp, const pf_box& box, void* dst) override;
  void read_box_q(int comp, const pf_box& box, int bits, void* dst, double* scale) override;
  void write_field(int comp, const void* src) override;
  void dft_read(int mon, int slot, double* dst) override;
  void dft_flush() override;
  void info(pf_info* out) const override;
  void field_view(int comp, pf_view* out) override;
  void dft_view(int mon, pf_view* out) override;
  void flux_read(double* dst) override;
  void timing_enable(int every) override;
  void checkpoint_enable(int interval) override;
  void adjoint(int n_cot, const pf_cotangent* cot, double* grad) override;
  // batched 2-D stepping support
  // Nodes of this simulation that need the batched post pass.
  struct PostNodes {
    std::vector<std::pair<int, int>> ade;    // (cell, (medium << 26) | index)
    std::vector<std::pair<int, int>> esrc;   // (cell, E-side source group)
    struct Dft {
      int cell, mon, slot, comp;
      int64_t r;
    };
    std::vector<Dft> dft;
  };
  void describe(pfhip::SimDev<real>& s, int pol, PostNodes& post);
  void reset(hipStream_t s);
  void set_level(int p) { cur_ = p; }
  void write_ce(const void* src, hipStream_t s) {
    HIP_CHECK(hipMemcpyAsync(const_cast<real*>(args_.ce), src, (size_t)ncell() * sizeof(real),
                             hipMemcpyHostToDevice, s));
    if (zpad_ || ypad_) {
      HIP_CHECK(hipStreamSynchronize(s));
      const std::vector<real> p =
          zpad_ ? pad_rows(static_cast<const real*>(src), (int)shape[0], (int)shape[1], (int)shape[2],
                           zrow_, zpad_)
                : pad_rows(static_cast<const real*>(src), (int)shape[0], 1, (int)shape[1], yrow_, ypad_);
      copy_sync(ce_f_, p.data(), p.size() * sizeof(real), hipMemcpyHostToDevice, s);
    }
  }
  hipStream_t stream() const { return stream_; }
  int64_t ncell() const { return shape[0] * shape[1] * shape[2]; }
  void timing_read(double* ms, int64_t* launches) override;
  // batched 3-D stepping support: the batch launches the march kernels for
  // all its simulations (tiles from march_tile_list, arguments from
  // march_args) on the stream it lends (adopt_stream).
  void check_march_batch() const;
  std::vector<pfhip::MarchTile> march_tile_list(int nw, int xchunk, bool cpml) const;
  const pfhip::MarchArgs<real>& march_args(int p) const { return ma_[p]; }
  int level() const { return cur_; }
  int march_dist() const { return march_dist_; }
  int zpad() const { return zpad_; }
  void after_march(int64_t step);
  void post_step(int64_t step, const pfhip::AdeJob<real>* tab);
  // Device tables of the ADE media for one launch of all media per step,
  // indexed [field level read][polarisation level] (the two-step march
  // swaps the polarisation buffers after every pair, so a single step can
  // run at either P level). Built on first use of each pair of levels when
  // there are several media; their cell lists are disjoint, each cell
  // having one owner.
  pfhip::AdeJob<real>* ade_tab1_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int ade_n_jobs1_ = 0;
  int64_t ade_max_n1_ = 0;
  std::vector<int64_t> ade_job_n_;   // nodes per job of those tables
  // ADE media, E-side sources and DFT samples after the march in one launch
  // (k_post_step): their E nodes are disjoint
  bool post_ok_ = false;
  void adopt_stream(hipStream_t s);
  // The batched ADE and source passes of a 3-D batch: this simulation's
  // jobs for a step that reads level p; finish_step then flips the level
  // and captures the DFT samples of that step.
  void ade_jobs(int p, std::vector<pfhip::AdeJob<real>>& out) const;
  bool src_job(int p, pfhip::SrcJob<real>& out) const;
  void finish_step(int64_t step);
  // Batched DFT capture: this simulation's monitor jobs for samples written
  // from level p from step0 on; finish_step_dft then flips the level and
  // advances the rings' fill (folding a full ring) for a step whose samples
  // the batched launch wrote.
  void dft_jobs(int p, int64_t step0, std::vector<pfhip::DftJob<real>>& out) const;
  void finish_step_dft(int64_t step);
  // True when the sample of `step` fills a DFT ring (finish_step_dft folds it)
  bool dft_ring_fills(int64_t step) const {
    for (const Dft& m : dft_)
      if (step % m.interval == 0 && m.count + 1 == m.K) return true;
    return false;
  }

 private:
  struct Ade {
    int64_t n;
    int n_poles;
    int64_t* cells;
    real* abc;
    real inv_eps;
    real* En;
    real* Pcur;
    real* Pprev;
    real* Pcur1 = nullptr;   // second level of the two-step march
    real* Pprev1 = nullptr;
    bool is_box = false;     // cells are the box [lo, hi) in x-major order
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  };
  struct Dft {
    int n_comp;
    int32_t comps[6];
    int32_t* comps_dev;
    int32_t* is_h_dev;
    BoxDev box;
    int64_t nbox;
    int n_freq;
    int interval;
    int64_t n_samples;
    double* phase_e;
    double* phase_h;
    double dt;
    double* acc;     // [n_comp][n_freq][nbox][2]
    real* ring;      // [K][n_comp][nbox]
    int K;
    int count;       // buffered samples
    int64_t sample0; // sample index of ring slot 0
  };
  struct Flux {
    int axis;
    int64_t index;
    double dA, dt;
    int n_part;
    double* partial;
    double* total;
  };

  void step_once(int64_t step, bool monitors = true);
  void launch_march_lists(const pfhip::MarchArgs<real>& base);
  void save_state(real* dst);
  void capture_dft(int64_t step);
  bool dft_take(Dft& m, int64_t step);
  void fold(Dft& m);
  real* staging(size_t elems);
  const real* gather(int comp, const pf_box& box, int64_t* nbox);
  template <typename F>
  void timed(int cls, F&& launch);
  void collect_timing();

  DeviceArena arena_;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = true;       // false once a batch lends its stream
  bool exact_;
  int64_t n_steps_;
  StepArgs<real> args_;         // generic in-place kernels (level 0)
  real* f_[2][6] = {};           // field arrays per time level (level 1 only when fused)
  real** fields_dev_[2] = {};    // device copies of f_[level]
  int cur_ = 0;                  // level holding the current fields
  bool fused_ = false;           // fused ping-pong step in use
  pfhip::FusedArgs<real> fa_[2]; // fused-step arguments reading level p
  bool slab_ok_ = true;          // CPML nodes form [0, lo) U [hi, n) on every axis
  pfhip::MarchArgs<real> ma_[2]; // march-kernel arguments reading level p
  // Field levels of the march use padded z rows (pfhip::zslot) when zpad_:
  // rows of zrow_ slots, fcell_ elements per component, ce in ce_f_.
  int zpad_ = 0;                 // nodes per padded block: 64 (dense rows padded at the
                                 // end, the 64-lane march), 62, or 60 with the two-step
                                 // march; 0: dense rows with the 62-node march
  // 2-D batch members: padded y rows (ypad_ columns per 64-slot block) of yrow_ slots
  int ypad_ = 0;
  int yrow_ = 0;
  int zrow_ = 0;
  int64_t fcell_ = 0;
  real* ce_f_ = nullptr;
  real* lvl_[2] = {nullptr, nullptr};   // six components per time level
  real* psi_side_[2] = {nullptr, nullptr};  // E-side psi, H-side psi (level 0)
  real* psi_h_lvl1_ = nullptr;           // H-side psi, level 1
  size_t psi_len_h_ = 0;
  real* coef_ = nullptr;
  // March tile lists: [0] psi-free, [1] CPML on several axes (or any axis
  // with padded rows); with dense rows [2] z-only, [3] y-only, [4] x-only CPML
  // tiles, each launched with the kernel variant of its axis set. (Lists of
  // y+z and x+z tiles with their own variants, 80 VGPRs, measured slower than
  // running them on the several-axis variant.)
  static constexpr int NLIST = 5;
  static constexpr int list_mask_[NLIST] = {0, 7, 4, 2, 1};
  int* tiles_[NLIST] = {};
  int ntiles_[NLIST] = {};
  int64_t tile_cells_[NLIST] = {};       // owned cells of each list
  // Lists launched together with the psi-free list (bit q: list q; lists 2
  // and 3, z- and y-only CPML): one launch over the union in spatial order,
  // so CPML tiles and their psi-free neighbours share the apron lines.
  int mix_mask_ = 0;
  // lists launched on the side stream (bit q: list q): all-axis, y- and
  // x-only; with the mixed launch the all-axis list only (the x-only list
  // follows the mixed launch on the main stream: at 512^3 float32 the side
  // stream's two lists outlasted the mixed launch by 170 us, one by 80 us)
  int side_mask_ = -1;
  bool even_x_ = false;           // x chunks of nearly equal length (64-lane march)
  int* tiles_mix_ = nullptr;
  int ntiles_mix_ = 0;
  int march_nw_ = pfhip::MARCH_NW;
  int pml_nw_ = pfhip::MARCH_NW;  // wavefronts per workgroup of the CPML-tile launch
  int march_dist_ = 2;            // planes of prefetch in the march kernel
  bool band_order_ = false;       // tile order: all XCDs on one x-chunk at a time
  bool split_ = false;           // separate psi-free launch for tiles off the CPML slabs
  hipStream_t side_ = nullptr;   // second stream for the CPML-tile launch of a split step
  hipEvent_t fork_ev_[2] = {nullptr, nullptr};
  // two-step march (temporal blocking): one launch advances two steps
  bool pair_ = false;
  int pair_nw_ = 10;
  real* psi_e_lvl_[2] = {nullptr, nullptr};  // E-side psi, two levels
  int psi_e_cur_ = 0;                        // level holding the current E-side psi
  int* tiles2_[2] = {nullptr, nullptr};      // pair tiles: [0] psi-free, [1] with CPML
  int ntiles2_[2] = {0, 0};
  pfhip::MarchArgs2<real> m2_;               // launch-invariant part of the pair arguments
  pfhip::March2Ade<real>* ade_dev_[2] = {nullptr, nullptr};  // media, reading P level 0 / 1
  pfhip::March2Dft<real>* dft_dev_ = nullptr;
  int p_lvl_ = 0;                            // P level holding the current polarisation
  void setup_pair(int xchunk);
  void build_pair_tiles(int xchunk);
  void set_psi_e(int lvl);
  void step_pair(int64_t step);
  bool src_e_sorted_ = true;      // E-side source groups strictly ordered by (component, cell)
  int src_e_beg_[4] = {0, 0, 0, 0};   // first E-side group of each component
  int src_e_ilo_ = 1, src_e_ihi_ = 0; // x-plane range of the E-side sources
  // checkpoints for the adjoint (generic layout: level 0 fields + both psi sides)
  int ckpt_interval_ = 0;
  std::vector<real*> ckpts_;
  size_t psi_len_e_ = 0;
  int64_t steps_done_ = 0;
  void build_march_tiles(int xchunk);
  void check_march_tiles(const std::vector<pfhip::MarchTile>& t, int nw) const;
  int slab_e_[3][2], slab_h_[3][2];
  SourceGroupsDev src_h_, src_e_;
  std::vector<Ade> ade_;
  std::vector<Dft> dft_;
  std::vector<Flux> flux_;
  real* stage_ = nullptr;
  size_t stage_elems_ = 0;
  void* qbuf_ = nullptr;          // quantised snapshot (1 or 2 bytes per value)
  size_t qbuf_bytes_ = 0;
  double* qpart_ = nullptr;       // partial maxima + the scale of read_box_q
  int32_t* comp_index_dev_ = nullptr;  // {0, 1, 2, 3, 4, 5}
  // event timing
  struct Timed {
    hipEvent_t a, b;
    int cls;
    bool shared_start;   // a is the previous entry's b
  };
  bool timing_ = false;
  // time the step groups (march, aux) of every timing_every_-th step only;
  // other classes (DFT folds, pairs) in every step
  int timing_every_ = 1;
  bool skip_step_groups_ = false;
  std::vector<Timed> pending_;
  std::vector<hipEvent_t> pool_;
  int open_cls_ = -1;            // kernel class of the timed group being issued
  hipEvent_t open_ev_ = nullptr; // its start event
  // close event of the last top-level group while nothing has been issued on
  // the stream after it: the next group starts at that event instead of
  // recording one of its own (two events per step fewer)
  hipEvent_t chain_ev_ = nullptr;
  hipEvent_t record_event();
  double t_ms_[PF_NKCLASS] = {};
  int64_t t_count_[PF_NKCLASS] = {};
};

template <typename real>
Stepper<real>::Stepper(const pf_desc& d, int zpad_default) {
  for (int a = 0; a < 3; ++a) {
    shape[a] = d.shape[a];
    if (d.shape[a] < 1) throw ArgError{"shape must be positive"};
  }
  if (d.shape[0] * d.shape[1] * d.shape[2] >= (int64_t(1) << 31))
    throw ArgError{"grids of 2^31 nodes or more are not supported"};
  if (d.shape[1] * d.shape[2] >= (int64_t(1) << 31) || d.shape[0] > 65535 * 16)
    throw ArgError{"grid extent beyond the launch limits"};
  if (!d.ce) throw ArgError{"ce is required"};
  for (int a = 0; a < 3; ++a)
    if (!d.b_e[a] || !d.c_e[a] || !d.b_h[a] || !d.c_h[a])
      throw ArgError{"CPML coefficient arrays are required on every axis"};
  if (d.n_steps < 0 || d.n_src < 0 || d.n_wave < 0 || d.n_ade < 0 || d.n_dft < 0 || d.n_flux < 0)
    throw ArgError{"negative table length"};
  if (d.n_src > 0 && (!d.src_comp || !d.src_cell || !d.src_amp || !d.src_wave || !d.wave))
    throw ArgError{"source tables missing"};

  HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  exact_ = d.exact != 0;
  n_steps_ = d.n_steps;
  const int nx = (int)d.shape[0], ny = (int)d.shape[1], nz = (int)d.shape[2];
  const int64_t ncell = (int64_t)nx * ny * nz;
  std::memset(&args_, 0, sizeof(args_));
  args_.nx = nx;
  args_.ny = ny;
  args_.nz = nz;
  args_.nyz = (int64_t)ny * nz;
  zrow_ = nz;
  fcell_ = ncell;
  for (int a = 0; a < 3; ++a) {
    args_.d[a] = (real)d.cell[a];
    args_.inv_d[a] = (real)(1.0 / d.cell[a]);
  }
  args_.ch = (real)d.ch;
  // one buffer per time level: the six components back to back
  lvl_[0] = arena_.zeros<real>(6 * (size_t)ncell, stream_);
  for (int c = 0; c < 6; ++c) args_.f[c] = lvl_[0] + (size_t)c * ncell;
  args_.ce = arena_.upload(static_cast<const real*>(d.ce), ncell, stream_);
  if ((d.ce_y == nullptr) != (d.ce_z == nullptr))
    throw ArgError{"ce_y and ce_z are given together (subpixel smoothing) or not at all"};
  if (d.ce_y) {
    args_.ce_y = arena_.upload(static_cast<const real*>(d.ce_y), ncell, stream_);
    args_.ce_z = arena_.upload(static_cast<const real*>(d.ce_z), ncell, stream_);
  }

  // CPML tables: coefficients rounded to the field dtype, in one table laid
  // out [axis][b_e, c_e, b_h, c_h][node]; compact maps per axis.
  const int n_axis[3] = {nx, ny, nz};
  coef_ = static_cast<real*>(arena_.alloc(4 * (size_t)(nx + ny + nz) * sizeof(real)));
  size_t coef_off = 0;
  for (int a = 0; a < 3; ++a) {
    const int n = n_axis[a];
    std::vector<real> be(n), ce(n), bh(n), ch(n);
    std::vector<int> me(n, -1), mh(n, -1);
    int pe = 0, ph = 0;
    for (int i = 0; i < n; ++i) {
      be[i] = (real)d.b_e[a][i];
      ce[i] = (real)d.c_e[a][i];
      bh[i] = (real)d.b_h[a][i];
      ch[i] = (real)d.c_h[a][i];
      if (d.c_e[a][i] != 0.0) me[i] = pe++;
      if (d.c_h[a][i] != 0.0) mh[i] = ph++;
    }
    real* tbl = coef_ + coef_off;
    copy_sync(tbl, be.data(), n * sizeof(real), hipMemcpyHostToDevice, stream_);
    copy_sync(tbl + n, ce.data(), n * sizeof(real), hipMemcpyHostToDevice, stream_);
    copy_sync(tbl + 2 * n, bh.data(), n * sizeof(real), hipMemcpyHostToDevice, stream_);
    copy_sync(tbl + 3 * n, ch.data(), n * sizeof(real), hipMemcpyHostToDevice, stream_);
    args_.b_e[a] = tbl;
    args_.c_e[a] = tbl + n;
    args_.b_h[a] = tbl + 2 * n;
    args_.c_h[a] = tbl + 3 * n;
    coef_off += 4 * (size_t)n;
    args_.map_e[a] = arena_.upload(me.data(), n, stream_);
    args_.map_h[a] = arena_.upload(mh.data(), n, stream_);
    args_.pe[a] = pe;
    args_.ph[a] = ph;
    slab_ok_ = slab_ok_ && two_slabs(me, slab_e_[a]) && two_slabs(mh, slab_h_[a]);
  }
  // psi: one buffer per side, its six slots back to back in slot order
  static const int slot_axis[12] = {1, 2, 2, 0, 0, 1, 1, 2, 2, 0, 0, 1};
  size_t psi_len[2] = {0, 0};
  for (int s = 0; s < pfhip::PSI_COUNT; ++s) {
    const int ax = slot_axis[s];
    const int pc = s < 6 ? args_.pe[ax] : args_.ph[ax];
    psi_len[s / 6] += (size_t)(ncell / n_axis[ax]) * pc;
  }
  psi_side_[0] = arena_.zeros<real>(psi_len[0], stream_);
  psi_side_[1] = arena_.zeros<real>(psi_len[1], stream_);
  {
    size_t off[2] = {0, 0};
    for (int s = 0; s < pfhip::PSI_COUNT; ++s) {
      const int ax = slot_axis[s];
      const int pc = s < 6 ? args_.pe[ax] : args_.ph[ax];
      args_.psi[s] = psi_side_[s / 6] ? psi_side_[s / 6] + off[s / 6] : nullptr;
      off[s / 6] += (size_t)(ncell / n_axis[ax]) * pc;
    }
  }
  for (int c = 0; c < 6; ++c) f_[0][c] = args_.f[c];
  psi_len_h_ = psi_len[1];
  psi_len_e_ = psi_len[0];
  fields_dev_[0] = arena_.upload(f_[0], 6, stream_);
  const int32_t comp_index[6] = {0, 1, 2, 3, 4, 5};
  comp_index_dev_ = arena_.upload(comp_index, 6, stream_);

  // Sources: consecutive equal (comp, cell) entries form one group.
  std::vector<int32_t> gcomp[2], gbeg[2];
  std::vector<int64_t> gcell[2];
  std::vector<double> amp[2];
  std::vector<int32_t> wrow[2];
  for (int s = 0; s < d.n_src; ++s) {
    const int comp = d.src_comp[s];
    if (comp < 0 || comp > 5) throw ArgError{"source component out of range"};
    if (d.src_cell[s] < 0 || d.src_cell[s] >= ncell) throw ArgError{"source cell out of range"};
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
  {
    // E-side groups for the two-step march: sorted by (component, cell),
    // per-component ranges and the x-planes they touch
    const std::vector<int32_t>& gc = gcomp[0];
    const std::vector<int64_t>& gx = gcell[0];
    for (size_t i = 1; i < gc.size(); ++i)
      if (gc[i] < gc[i - 1] || (gc[i] == gc[i - 1] && gx[i] <= gx[i - 1])) src_e_sorted_ = false;
    for (int c = 0; c < 4; ++c)
      src_e_beg_[c] = (int)(std::lower_bound(gc.begin(), gc.end(), c) - gc.begin());
    if (!src_e_sorted_) src_e_beg_[3] = (int)gc.size();
    for (int64_t cell : gx) {
      const int xi = (int)(cell / ((int64_t)ny * nz));
      src_e_ilo_ = src_e_ilo_ > src_e_ihi_ ? xi : std::min(src_e_ilo_, xi);
      src_e_ihi_ = std::max(src_e_ihi_, xi);
    }
  }
  double* wave = arena_.upload(d.wave, (size_t)d.n_wave * d.n_steps, stream_);
  SourceGroupsDev* groups[2] = {&src_e_, &src_h_};
  for (int side = 0; side < 2; ++side) {
    SourceGroupsDev& g = *groups[side];
    g.n_groups = (int32_t)gcomp[side].size();
    if (g.n_groups == 0) continue;
    gbeg[side].push_back((int32_t)amp[side].size());
    g.comp = arena_.upload(gcomp[side].data(), gcomp[side].size(), stream_);
    g.cell = arena_.upload(gcell[side].data(), gcell[side].size(), stream_);
    g.begin = arena_.upload(gbeg[side].data(), gbeg[side].size(), stream_);
    g.amp = arena_.upload(amp[side].data(), amp[side].size(), stream_);
    g.wave_row = arena_.upload(wrow[side].data(), wrow[side].size(), stream_);
    g.wave = wave;
  }

  // ADE media.
  for (int m = 0; m < d.n_ade; ++m) {
    const pf_ade_desc& a = d.ade[m];
    if (a.n_cells <= 0) continue;
    if (a.n_poles < 0 || (a.n_poles > 0 && !a.abc) || !a.cells) throw ArgError{"bad ADE table"};
    for (int64_t t = 0; t < a.n_cells; ++t)
      if (a.cells[t] < 0 || a.cells[t] >= ncell) throw ArgError{"ADE cell out of range"};
    Ade st;
    st.n = a.n_cells;
    st.n_poles = a.n_poles;
    st.cells = arena_.upload(a.cells, a.n_cells, stream_);
    std::vector<real> abc(3 * (size_t)a.n_poles);
    for (size_t i = 0; i < abc.size(); ++i) abc[i] = (real)a.abc[i];
    st.abc = arena_.upload(abc.data(), abc.size(), stream_);
    st.inv_eps = (real)a.inv_eps;
    st.En = arena_.zeros<real>(3 * a.n_cells, stream_);
    st.Pcur = arena_.zeros<real>(3 * (size_t)a.n_poles * a.n_cells, stream_);
    st.Pprev = arena_.zeros<real>(3 * (size_t)a.n_poles * a.n_cells, stream_);
    {
      // box-shaped media (cells enumerate [lo, hi) x-major) can run in the two-step march
      int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {-1, -1, -1};
      for (int64_t t = 0; t < a.n_cells; ++t) {
        const int64_t c = a.cells[t];
        const int64_t x[3] = {c / ((int64_t)ny * nz), (c / nz) % ny, c % nz};
        for (int q = 0; q < 3; ++q) {
          lo[q] = std::min(lo[q], x[q]);
          hi[q] = std::max(hi[q], x[q] + 1);
        }
      }
      const int64_t e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
      bool box = e[0] * e[1] * e[2] == a.n_cells;
      for (int64_t t = 0; box && t < a.n_cells; ++t) {
        const int64_t x = lo[0] + t / (e[1] * e[2]), y = lo[1] + (t / e[2]) % e[1], z = lo[2] + t % e[2];
        box = a.cells[t] == (x * ny + y) * nz + z;
      }
      st.is_box = box;
      for (int q = 0; q < 3; ++q) {
        st.lo[q] = (int)lo[q];
        st.hi[q] = (int)hi[q];
      }
    }
    ade_.push_back(st);
  }

  // DFT monitors: ring of K samples, folded into complex128 accumulators.
  for (int m = 0; m < d.n_dft; ++m) {
    const pf_dft_desc& s = d.dft[m];
    Dft st;
    if (s.n_comp < 1 || s.n_comp > 6 || s.n_freq < 1 || s.interval < 1 || s.n_samples < 0)
      throw ArgError{"bad DFT monitor description"};
    st.n_comp = s.n_comp;
    int32_t is_h[6];
    for (int c = 0; c < s.n_comp; ++c) {
      if (s.comps[c] < 0 || s.comps[c] > 5) throw ArgError{"DFT component out of range"};
      st.comps[c] = s.comps[c];
      is_h[c] = s.comps[c] >= 3;
    }
    st.comps_dev = arena_.upload(st.comps, s.n_comp, stream_);
    st.is_h_dev = arena_.upload(is_h, s.n_comp, stream_);
    st.box = check_box(s.box, d.shape);
    st.nbox = st.box.count[0] * st.box.count[1] * st.box.count[2];
    st.n_freq = s.n_freq;
    st.interval = s.interval;
    st.n_samples = s.n_samples;
    st.phase_e = arena_.upload(s.phase_e, (size_t)s.n_samples * s.n_freq * 2, stream_);
    st.phase_h = arena_.upload(s.phase_h, (size_t)s.n_samples * s.n_freq * 2, stream_);
    st.dt = s.dt;
    st.acc = arena_.zeros<double>((size_t)s.n_comp * s.n_freq * st.nbox * 2, stream_);
    const int64_t sample_bytes = (int64_t)s.n_comp * st.nbox * (int64_t)sizeof(real);
    // ring of up to 128 samples (1 GB per monitor at most): each fold rewrites
    // the complex128 accumulators, so fewer folds save that traffic
    const int64_t budget = 1ll << 30;
    st.K = (int)std::max<int64_t>(1, std::min<int64_t>(128, budget / std::max<int64_t>(sample_bytes, 1)));
    // a smaller ring on request (tests of the fold schedule)
    if (const char* env_ring = std::getenv("PHOTONFDTD_HIP_DFT_RING"))
      st.K = std::max(1, std::min(st.K, std::atoi(env_ring)));
    st.ring = static_cast<real*>(arena_.alloc((size_t)st.K * sample_bytes));
    st.count = 0;
    st.sample0 = 0;
    dft_.push_back(st);
  }

  {
    // One launch for the ADE media, E-side sources and DFT samples after the
    // march needs disjoint E nodes: no source on a dispersive cell and no
    // sampled E node that either of them writes (PHOTONFDTD_HIP_POST=0: one
    // launch each). At 512^3 float32 (two media, one dipole, two planes of
    // four components) 18 us per step against 33 us for four launches.
    const char* env_post = std::getenv("PHOTONFDTD_HIP_POST");
    int n_media = 0;
    std::vector<int64_t> disp;
    for (int m = 0; m < d.n_ade; ++m)
      if (d.ade[m].n_cells > 0 && d.ade[m].n_poles > 0) {
        ++n_media;
        disp.insert(disp.end(), d.ade[m].cells, d.ade[m].cells + d.ade[m].n_cells);
      }
    std::sort(disp.begin(), disp.end());
    bool ok = !(env_post && env_post[0] == '0') && n_media <= pfhip::POST_MAX_ADE &&
              (int)dft_.size() <= pfhip::POST_MAX_DFT;
    for (int64_t c : gcell[0])
      if (ok && std::binary_search(disp.begin(), disp.end(), c)) ok = false;
    auto in_box = [&](const BoxDev& b, int64_t c) {
      const int64_t x[3] = {c / ((int64_t)ny * nz), (c / nz) % ny, c % nz};
      for (int q = 0; q < 3; ++q) {
        const int64_t u = x[q] - b.start[q];
        if (u < 0 || u >= b.count[q] * b.stride[q] || u % b.stride[q] != 0) return false;
      }
      return true;
    };
    for (const Dft& m : dft_) {
      bool comp_e[3] = {false, false, false};
      for (int c = 0; c < m.n_comp; ++c)
        if (m.comps[c] < 3) comp_e[m.comps[c]] = true;
      if (!(comp_e[0] || comp_e[1] || comp_e[2])) continue;
      for (size_t t = 0; ok && t < disp.size(); ++t)
        if (in_box(m.box, disp[t])) ok = false;
      for (size_t g = 0; ok && g < gcell[0].size(); ++g)
        if (comp_e[gcomp[0][g]] && in_box(m.box, gcell[0][g])) ok = false;
    }
    post_ok_ = ok;
  }

  for (int m = 0; m < d.n_flux; ++m) {
    const pf_flux_desc& s = d.flux[m];
    if (s.axis < 0 || s.axis > 2 || s.index < 0 || s.index >= d.shape[s.axis])
      throw ArgError{"flux plane outside the grid"};
    Flux st;
    st.axis = s.axis;
    st.index = s.index;
    st.dA = s.dA;
    st.dt = s.dt;
    const int64_t plane = ncell / d.shape[s.axis];
    st.n_part = (int)std::min<int64_t>(512, (plane + 255) / 256);
    st.partial = arena_.zeros<double>(st.n_part, stream_);
    st.total = arena_.zeros<double>(1, stream_);
    flux_.push_back(st);
  }

  // Fused ping-pong step: 3-D grids without H-side sources (an H source
  // must land between the H and E halves of the step).
  const char* env_fused = std::getenv("PHOTONFDTD_HIP_FUSED");
  // (per-component coefficients of subpixel smoothing run on the generic kernels)
  const bool want_fused = !(env_fused && env_fused[0] == '0') && !(d.flags & PF_FLAG_GENERIC) &&
                          d.ce_y == nullptr;
  if (want_fused && pfhip::fused_supported<real>(args_) && src_h_.n_groups == 0) {
    fused_ = true;
    const char* env_var = std::getenv("PHOTONFDTD_HIP_VARIANT");
    int variant = env_var ? std::atoi(env_var) : 2;
    if (variant == 2 && !slab_ok_) variant = 1;
    // Padded z rows for the marches (whole-line row stores).
    const char* env_zpad = std::getenv("PHOTONFDTD_HIP_ZPAD");
    // 62-node blocks for the single-step march's tiles; 60 when the
    // two-step march (two-node apron) shares the levels
    const char* env_pair = std::getenv("PHOTONFDTD_HIP_PAIR");
    // 64: rows padded at the end only, the 64-lane march with one launch
    // per CPML axis set (the default from 2^25 nodes: its psi-free tiles
    // pay for the extra launches; 1.74 against 2.01 ms per step at 512^3);
    // 62: 62 nodes per 64-slot block, two launches (the default below, where
    // most tiles touch a slab: 0.082 against 0.113 ms at 128^3); 0: dense
    // rows with the 62-node march
    const int zdef = zpad_default ? zpad_default
                     : (d.flags & PF_FLAG_DENSE64) && nz % 64 == 0 ? pfhip::ZPAD_BLOCK
                     : ncell >= (int64_t(1) << 25) ? pfhip::ZPAD_BLOCK : 62;
    const int zenv = env_zpad ? std::atoi(env_zpad) : -1;
    zpad_ = variant == 2 && zenv != 0
                ? (env_pair && env_pair[0] == '1'
                       ? pfhip::MARCH2_OWNED_Z
                       : zenv == 62 || zenv == pfhip::ZPAD_BLOCK ? zenv : zdef)
                : 0;
    if (zpad_ == pfhip::ZPAD_BLOCK) {
      // the 64-lane march addresses three field components, and each psi
      // side, through one buffer resource with 32-bit offsets below 2^31
      const double lim = 2147483648.0;
      const double comp = (double)nx * ny * pfhip::zpad_row(nz, zpad_) * sizeof(real);
      if (3.0 * comp >= lim || (double)psi_len_e_ * sizeof(real) >= lim ||
          (double)psi_len_h_ * sizeof(real) >= lim)
        zpad_ = 62;
    }
    if (zpad_) {
      zrow_ = pfhip::zpad_row(nz, zpad_);
      fcell_ = (int64_t)nx * ny * zrow_;
      arena_.free(lvl_[0]);
      lvl_[0] = arena_.zeros<real>(6 * (size_t)fcell_, stream_);
      for (int c = 0; c < 6; ++c) {
        f_[0][c] = lvl_[0] + (size_t)c * fcell_;
        args_.f[c] = nullptr;   // the generic kernels never run on this stepper
      }
      fields_dev_[0] = arena_.upload(f_[0], 6, stream_);
      const std::vector<real> cep = pad_rows(static_cast<const real*>(d.ce), nx, ny, nz, zrow_, zpad_);
      ce_f_ = arena_.upload(cep.data(), cep.size(), stream_);
      // cell indices of the ADE media and the E-side sources in padded rows
      size_t ia = 0;
      for (int m = 0; m < d.n_ade; ++m) {
        const pf_ade_desc& ad = d.ade[m];
        if (ad.n_cells <= 0) continue;
        std::vector<int64_t> cells(ad.cells, ad.cells + ad.n_cells);
        for (int64_t& c : cells) c = pad_cell(c, nz, zrow_, zpad_);
        ade_[ia++].cells = arena_.upload(cells.data(), cells.size(), stream_);
      }
      if (src_e_.n_groups > 0) {
        std::vector<int64_t> cells = gcell[0];
        for (int64_t& c : cells) c = pad_cell(c, nz, zrow_, zpad_);
        src_e_.cell = arena_.upload(cells.data(), cells.size(), stream_);
      }
    }
    lvl_[1] = arena_.zeros<real>(6 * (size_t)fcell_, stream_);
    for (int c = 0; c < 6; ++c) f_[1][c] = lvl_[1] + (size_t)c * fcell_;
    fields_dev_[1] = arena_.upload(f_[1], 6, stream_);
    psi_h_lvl1_ = arena_.zeros<real>(psi_len_h_, stream_);
    real* psi_h1[6];
    for (int s = 0; s < 6; ++s)
      psi_h1[s] = psi_h_lvl1_ ? psi_h_lvl1_ + (args_.psi[6 + s] - args_.psi[6]) : nullptr;
    // planes per march tile: with dense rows at most 36 (x ranges cut into
    // chunks of nearly equal length), measured fastest at 512^3 float32
    // with the mixed launch (32-48 planes within 2 %)
    const char* env_xc = std::getenv("PHOTONFDTD_HIP_XCHUNK");
    const int xchunk = env_xc ? std::max(1, std::atoi(env_xc)) : zpad_ == pfhip::ZPAD_BLOCK ? 36 : 128;
    for (int p = 0; p < 2; ++p) {
      pfhip::FusedArgs<real>& fa = fa_[p];
      std::memset(&fa, 0, sizeof(fa));
      for (int c = 0; c < 3; ++c) {
        fa.Ein[c] = f_[p][c];
        fa.Eout[c] = f_[p ^ 1][c];
        fa.Hin[c] = f_[p][3 + c];
        fa.Hout[c] = f_[p ^ 1][3 + c];
      }
      fa.ce = args_.ce;
      for (int s = 0; s < 6; ++s) {
        fa.psiE[s] = args_.psi[s];
        real* lvl0 = args_.psi[6 + s];
        fa.psiHin[s] = p == 0 ? lvl0 : psi_h1[s];
        fa.psiHout[s] = p == 0 ? psi_h1[s] : lvl0;
      }
      for (int a = 0; a < 3; ++a) {
        fa.b_e[a] = args_.b_e[a];
        fa.c_e[a] = args_.c_e[a];
        fa.b_h[a] = args_.b_h[a];
        fa.c_h[a] = args_.c_h[a];
        fa.map_e[a] = args_.map_e[a];
        fa.map_h[a] = args_.map_h[a];
        fa.pe[a] = args_.pe[a];
        fa.ph[a] = args_.ph[a];
        fa.d[a] = args_.d[a];
        fa.inv_d[a] = args_.inv_d[a];
      }
      fa.nx = nx;
      fa.ny = ny;
      fa.nz = nz;
      fa.nyz = args_.nyz;
      fa.ch = args_.ch;
      fa.xchunk = xchunk;
      fa.variant = variant;
      pfhip::MarchArgs<real>& ma = ma_[p];
      std::memset(&ma, 0, sizeof(ma));
      ma.Lin = lvl_[p];
      ma.Lout = lvl_[p ^ 1];
      ma.ce = zpad_ ? ce_f_ : args_.ce;
      ma.zrow = zrow_;
      ma.zpad = zpad_;
      ma.psiE = psi_side_[0];
      ma.psiHin = p == 0 ? psi_side_[1] : psi_h_lvl1_;
      ma.psiHout = p == 0 ? psi_h_lvl1_ : psi_side_[1];
      ma.coef = coef_;
      ma.nx = nx;
      ma.ny = ny;
      ma.nz = nz;
      for (int a = 0; a < 3; ++a) {
        ma.pe[a] = args_.pe[a];
        ma.ph[a] = args_.ph[a];
        ma.d[a] = args_.d[a];
        ma.inv_d[a] = args_.inv_d[a];
        for (int q = 0; q < 2; ++q) {
          ma.slab_e[a][q] = slab_e_[a][q];
          ma.slab_h[a][q] = slab_h_[a][q];
        }
      }
      ma.ch = args_.ch;
      ma.xchunk = std::min(xchunk, pfhip::MARCH_MAX_XCHUNK);
      for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 2; ++q) {
          fa.slab_e[a][q] = slab_e_[a][q];
          fa.slab_h[a][q] = slab_h_[a][q];
        }
    }
  }

  // 12-wave workgroups by default with padded rows (measured fastest at
  // 512^3 float32 there), 10 with dense rows
  const char* env_nw = std::getenv("PHOTONFDTD_HIP_NW");
  if (zpad_) march_nw_ = 12;
  if (env_nw && (std::atoi(env_nw) == 16 || std::atoi(env_nw) == 12 || std::atoi(env_nw) == 10))
    march_nw_ = std::atoi(env_nw);
  const char* env_dist = std::getenv("PHOTONFDTD_HIP_PREFETCH");
  if (env_dist && std::atoi(env_dist) == 1) march_dist_ = 1;
  const char* env_order = std::getenv("PHOTONFDTD_HIP_ORDER");
  band_order_ = env_order && env_order[0] == '1';
  // Separate psi-free launch for the tiles off the CPML slabs: the default
  // with padded rows (the psi-free kernel, 72 VGPRs, runs two workgroups
  // per CU there).
  const char* env_split = std::getenv("PHOTONFDTD_HIP_SPLIT");
  split_ = env_split ? env_split[0] == '1' : zpad_ != 0;
  pml_nw_ = march_nw_;
  const char* env_pnw = std::getenv("PHOTONFDTD_HIP_PML_NW");
  if (split_ && env_pnw && (std::atoi(env_pnw) == 10 || std::atoi(env_pnw) == 12 || std::atoi(env_pnw) == 16))
    pml_nw_ = std::atoi(env_pnw);
  const char* env_smask = std::getenv("PHOTONFDTD_HIP_SIDE");
  side_mask_ = env_smask ? std::atoi(env_smask) & 0b11110 : -1;
  const char* env_side = std::getenv("PHOTONFDTD_HIP_OVERLAP");
  if (fused_ && split_ && !(env_side && env_side[0] == '0')) {
    HIP_CHECK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    for (hipEvent_t& e : fork_ev_) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (fused_ && fa_[0].variant == 2) build_march_tiles(std::min(fa_[0].xchunk, pfhip::MARCH_MAX_XCHUNK));
  if (fused_ && fa_[0].variant == 2) setup_pair(std::min(fa_[0].xchunk, pfhip::MARCH_MAX_XCHUNK));
  HIP_CHECK(hipStreamSynchronize(stream_));
}

// Tiles of the march kernel: tiles whose every node (including the H apron
// row/lane and the plane before the chunk) lies outside the CPML slabs go to
// the psi-free kernel, the rest to the full kernel.
template <typename real>
void Stepper<real>::build_march_tiles(int xchunk) {
  const int n[3] = {args_.nx, args_.ny, args_.nz};
  const int len[3] = {xchunk, march_nw_ - 2, 62};
  const char* env_even = std::getenv("PHOTONFDTD_HIP_EVEN_X");
  even_x_ = !(env_even && env_even[0] == '0');
  std::vector<Seg> seg[3];
  for (int a = 0; a < 3; ++a) {
    const int b0 = std::max(slab_e_[a][0], slab_h_[a][0] + 1);
    const int b1 = std::min(slab_e_[a][1], slab_h_[a][1]);
    if (a > 0 && zpad_ == pfhip::ZPAD_BLOCK && split_) {
      seg[a] = split_aligned(n[a], b0, b1, a == 2 ? pfhip::ZPAD_BLOCK : len[a]);
    } else if (a == 2 && zpad_ == pfhip::ZPAD_BLOCK) {
      seg[2] = split_aligned(n[2], 0, n[2], pfhip::ZPAD_BLOCK);
    } else if (a == 2 && zpad_) {
      // padded rows: z tiles on the fixed grid of zpad_ nodes (each owns
      // one 64-slot block); a tile is psi-free when it lies inside [b0, b1)
      for (int z = 0; z < n[2]; z += zpad_) {
        const int c = std::min(zpad_, n[2] - z);
        seg[2].push_back({z, c, z >= b0 && z + c <= b1});
      }
    } else if (a == 0 && zpad_ == pfhip::ZPAD_BLOCK && split_ && even_x_) {
      // x chunks of nearly equal length (at most xchunk planes) in each of
      // the CPML and bulk ranges: no short last chunk re-reading its two
      // boundary planes for a few planes of work
      seg[0] = split_even(n[0], b0, b1, len[0]);
    } else {
      seg[a] = split_ ? split_axis(n[a], b0, b1, len[a], a > 0) : split_axis(n[a], 0, 0, len[a], false);
    }
  }
  std::vector<pfhip::MarchTile> lists[NLIST];   // [0] psi-free, [1] several CPML axes, ...
  std::vector<pfhip::MarchTile>& pml = lists[1];
  for (const Seg& sx : seg[0])
    for (const Seg& sy : seg[1])
      for (const Seg& sz : seg[2]) {
        pfhip::MarchTile t{sx.start, sx.start + sx.count, sy.start, sy.count, sz.start, sz.count, 0, 0};
        const int mask = (sx.bulk ? 0 : 1) | (sy.bulk ? 0 : 2) | (sz.bulk ? 0 : 4);
        if (!split_) {
          pml.push_back(t);
        } else if (zpad_ == pfhip::ZPAD_BLOCK) {
          int q = 0;
          while (q < NLIST && list_mask_[q] != mask) ++q;
          lists[q < NLIST ? q : 1].push_back(t);
        } else {
          lists[mask ? 1 : 0].push_back(t);
        }
      }
  if (split_) {
    // CPML tiles in shorter x chunks: that launch holds one workgroup per
    // CU and few tiles, so it needs more of them to keep every CU busy
    // (and rows for its own workgroup height); 12 planes with the 64-lane
    // layout (0.4 % faster than 16 at 512^3 float32), 16 with padded rows
    const char* env_px = std::getenv("PHOTONFDTD_HIP_PML_XCHUNK");
    const int px = std::max(1, env_px ? std::atoi(env_px) : zpad_ == pfhip::ZPAD_BLOCK ? 12 : 16);
    const int py = pml_nw_ - 2;
    std::vector<pfhip::MarchTile> cut;
    for (const pfhip::MarchTile& t : pml)
      for (int x = t.i0; x < t.i1; x += px)
        for (int y = t.j0; y < t.j0 + t.jn; y += py) {
          pfhip::MarchTile u = t;
          u.i0 = x;
          u.i1 = std::min(t.i1, x + px);
          u.j0 = y;
          u.jn = std::min(py, t.j0 + t.jn - y);
          cut.push_back(u);
        }
    pml.swap(cut);
    // the single-axis CPML lists run at the CPML workgroup height too
    for (int q = 2; q < NLIST; ++q) {
      std::vector<pfhip::MarchTile> rows;
      for (const pfhip::MarchTile& t : lists[q])
        for (int y = t.j0; y < t.j0 + t.jn; y += py) {
          pfhip::MarchTile u = t;
          u.j0 = y;
          u.jn = std::min(py, t.j0 + t.jn - y);
          rows.push_back(u);
        }
      lists[q].swap(rows);
    }
  }
  // z-only and y-only CPML tiles in the psi-free launch (float32 default;
  // PHOTONFDTD_HIP_MIX = bit mask of lists 2, 3 and 4: z-, y- and x-only),
  // 12-wave workgroups. The float64 mixed kernel needs 141 VGPRs (one
  // workgroup per CU), so float64 keeps one launch per list.
  const char* env_mix = std::getenv("PHOTONFDTD_HIP_MIX");
  mix_mask_ = (env_mix ? std::atoi(env_mix) : sizeof(real) == 4 ? 0b1100 : 0) & 0b11100;
  if ((mix_mask_ & 0b10000) && !(mix_mask_ & 0b1000)) mix_mask_ &= ~0b10000;   // x only with y
  if (zpad_ != pfhip::ZPAD_BLOCK || !split_ || march_nw_ != 12 || pml_nw_ != 12 || lists[0].empty())
    mix_mask_ = 0;
  if (side_mask_ < 0) side_mask_ = mix_mask_ ? 0b00010 : 0b11010;
  if (mix_mask_) {
    std::vector<pfhip::MarchTile> m;
    for (int q = 0; q < NLIST; ++q) {
      if (q != 0 && !(mix_mask_ >> q & 1)) continue;
      check_march_tiles(lists[q], 12);
      for (pfhip::MarchTile t : lists[q]) {
        t.pad1 = list_mask_[q];
        m.push_back(t);
      }
    }
    std::sort(m.begin(), m.end(), [](const pfhip::MarchTile& u, const pfhip::MarchTile& v) {
      return std::tie(u.i0, u.j0, u.k0) < std::tie(v.i0, v.j0, v.k0);
    });
    std::vector<pfhip::MarchTile> ordered = xcd_order(m);
    ntiles_mix_ = (int)ordered.size();
    tiles_mix_ = reinterpret_cast<int*>(arena_.upload(ordered.data(), ordered.size(), stream_));
  }
  for (int q = 0; q < NLIST; ++q) {
    check_march_tiles(lists[q], q == 0 ? march_nw_ : pml_nw_);
    tile_cells_[q] = 0;
    for (const pfhip::MarchTile& t : lists[q]) tile_cells_[q] += (int64_t)(t.i1 - t.i0) * t.jn * t.kn;
    std::vector<pfhip::MarchTile> ordered =
        band_order_ ? xcd_band_order(lists[q]) : xcd_order(lists[q]);
    ntiles_[q] = (int)ordered.size();
    tiles_[q] = ordered.empty() ? nullptr
                                : reinterpret_cast<int*>(arena_.upload(ordered.data(), ordered.size(), stream_));
  }
}

// Two-step march set-up: the pair kernel needs box-shaped media, E-side
// sources sorted by (component, cell), at most MARCH2_MAX_DFT monitors and
// no flux planes (their per-step sums would need the intermediate level).
template <typename real>
void Stepper<real>::setup_pair(int xchunk) {
  const char* env = std::getenv("PHOTONFDTD_HIP_PAIR");
  if (!env || env[0] != '1') return;   // opt-in
  if (zpad_ && zpad_ != pfhip::MARCH2_OWNED_Z) return;   // padded blocks of the pair's tile width
  if (!flux_.empty() || !src_e_sorted_ || ade_.size() > (size_t)pfhip::MARCH2_MAX_ADE ||
      dft_.size() > (size_t)pfhip::MARCH2_MAX_DFT)
    return;
  for (const Ade& a : ade_)
    if (!a.is_box) return;
  for (const Dft& d : dft_)
    if (d.K > 255) return;   // ring slots are passed as 8-bit indices
  const char* env_nw = std::getenv("PHOTONFDTD_HIP_PAIR_NW");
  pair_nw_ = env_nw && std::atoi(env_nw) == 12 ? 12 : 10;
  psi_e_lvl_[0] = psi_side_[0];
  psi_e_lvl_[1] = psi_len_e_ ? arena_.zeros<real>(psi_len_e_, stream_) : nullptr;
  for (Ade& a : ade_) {
    const size_t n = 3 * (size_t)a.n_poles * a.n;
    if (!n) continue;
    a.Pcur1 = arena_.zeros<real>(n, stream_);
    a.Pprev1 = arena_.zeros<real>(n, stream_);
  }
  pfhip::MarchArgs2<real>& m = m2_;
  std::memset(&m, 0, sizeof(m));
  const pfhip::MarchArgs<real>& ma = ma_[0];
  m.ce = ma.ce;
  m.coef = ma.coef;
  m.nx = ma.nx;
  m.ny = ma.ny;
  m.nz = ma.nz;
  for (int a = 0; a < 3; ++a) {
    m.pe[a] = ma.pe[a];
    m.ph[a] = ma.ph[a];
    m.d[a] = ma.d[a];
    m.inv_d[a] = ma.inv_d[a];
    for (int q = 0; q < 2; ++q) {
      m.slab_e[a][q] = ma.slab_e[a][q];
      m.slab_h[a][q] = ma.slab_h[a][q];
    }
  }
  m.ch = ma.ch;
  m.zrow = ma.zrow;
  m.zpad = ma.zpad;
  m.n_steps = n_steps_;
  for (int c = 0; c < 4; ++c) m.src_beg[c] = src_e_beg_[c];
  m.src_i_lo = src_e_ilo_;
  m.src_i_hi = src_e_ihi_;
  m.src_cell = src_e_.cell;
  m.src_begin = src_e_.begin;
  m.src_amp = src_e_.amp;
  m.src_row = src_e_.wave_row;
  m.src_wave = src_e_.wave;
  m.n_ade = (int)ade_.size();
  if (!ade_.empty()) {
    for (int lv = 0; lv < 2; ++lv) {
      std::vector<pfhip::March2Ade<real>> v(ade_.size());
      for (size_t i = 0; i < ade_.size(); ++i) {
        const Ade& s = ade_[i];
        pfhip::March2Ade<real>& A = v[i];
        std::memset(&A, 0, sizeof(A));
        for (int q = 0; q < 3; ++q) {
          A.lo[q] = s.lo[q];
          A.hi[q] = s.hi[q];
        }
        A.n_poles = s.n_poles;
        A.n = s.n;
        A.abc = s.abc;
        A.inv_eps = s.inv_eps;
        real* cur[2] = {s.Pcur, s.Pcur1};
        real* prev[2] = {s.Pprev, s.Pprev1};
        A.Pcur_in = cur[lv];
        A.Pprev_in = prev[lv];
        A.Pcur_out = cur[lv ^ 1];
        A.Pprev_out = prev[lv ^ 1];
      }
      ade_dev_[lv] = arena_.upload(v.data(), v.size(), stream_);
    }
  }
  m.n_dft = (int)dft_.size();
  if (!dft_.empty()) {
    std::vector<pfhip::March2Dft<real>> v(dft_.size());
    for (size_t i = 0; i < dft_.size(); ++i) {
      pfhip::March2Dft<real>& D = v[i];
      std::memset(&D, 0, sizeof(D));
      D.n_comp = dft_[i].n_comp;
      for (int c = 0; c < dft_[i].n_comp; ++c) D.comps[c] = dft_[i].comps[c];
      for (int q = 0; q < 3; ++q) {
        D.start[q] = (int)dft_[i].box.start[q];
        D.count[q] = (int)dft_[i].box.count[q];
        D.stride[q] = (int)dft_[i].box.stride[q];
      }
      D.ring = dft_[i].ring;
      D.slot_elems = (int64_t)dft_[i].n_comp * dft_[i].nbox;
    }
    dft_dev_ = arena_.upload(v.data(), v.size(), stream_);
  }
  m.dft = dft_dev_;
  build_pair_tiles(xchunk);
  pair_ = true;
}

// Tiles of the pair kernel: jn <= NW-4 owned rows, kn <= 60 owned z nodes.
// A tile goes to the psi-free kernel when every node it computes (owned
// nodes plus two on each side, on every axis) lies off the CPML slabs.
template <typename real>
void Stepper<real>::build_pair_tiles(int xchunk) {
  const int n[3] = {args_.nx, args_.ny, args_.nz};
  const int len[3] = {xchunk, pair_nw_ - 4, pfhip::MARCH2_OWNED_Z};
  std::vector<Seg> seg[3];
  for (int a = 0; a < 3; ++a) {
    const int b0 = std::max(slab_e_[a][0], slab_h_[a][0]) + 2;
    const int b1 = std::min(slab_e_[a][1], slab_h_[a][1]) - 2;
    if (a == 2 && zpad_) {
      // padded rows: z tiles on the fixed 60-node grid (one 64-slot block each)
      for (int z = 0; z < n[2]; z += zpad_) {
        const int c = std::min(zpad_, n[2] - z);
        seg[2].push_back({z, c, z >= b0 && z + c <= b1});
      }
    } else {
      seg[a] = split_even(n[a], b0, b1, len[a]);
    }
  }
  std::vector<pfhip::MarchTile> bulk, pml;
  for (const Seg& sx : seg[0])
    for (const Seg& sy : seg[1])
      for (const Seg& sz : seg[2]) {
        // the kernel's launch shape assumes these bounds
        if (sx.count < 1 || sx.count > pfhip::MARCH_MAX_XCHUNK || sy.count < 1 ||
            sy.count > pair_nw_ - 4 || sz.count < 1 || sz.count > pfhip::MARCH2_OWNED_Z ||
            sx.start < 0 || sx.start + sx.count > n[0] || sy.start < 0 || sy.start + sy.count > n[1] ||
            sz.start < 0 || sz.start + sz.count > n[2] || (zpad_ && sz.start % zpad_ != 0))
          throw HipError{PF_E_STATE, "pair tile outside the kernel's bounds"};
        pfhip::MarchTile t{sx.start, sx.start + sx.count, sy.start, sy.count, sz.start, sz.count, 0, 0};
        (sx.bulk && sy.bulk && sz.bulk ? bulk : pml).push_back(t);
      }
  if (zpad_) {
    // CPML tiles in shorter x chunks (their launch runs beside the psi-free
    // one and needs more, shorter tiles to spread over the CUs)
    const char* env_px = std::getenv("PHOTONFDTD_HIP_PAIR_PML_XCHUNK");
    const int px = std::max(1, env_px ? std::atoi(env_px) : 32);
    std::vector<pfhip::MarchTile> cut;
    for (const pfhip::MarchTile& t : pml)
      for (int x = t.i0; x < t.i1; x += px) {
        pfhip::MarchTile u = t;
        u.i0 = x;
        u.i1 = std::min(t.i1, x + px);
        cut.push_back(u);
      }
    pml.swap(cut);
  }
  const std::vector<pfhip::MarchTile>* lists[2] = {&bulk, &pml};
  for (int q = 0; q < 2; ++q) {
    std::vector<pfhip::MarchTile> ordered = xcd_order(*lists[q]);
    ntiles2_[q] = (int)ordered.size();
    tiles2_[q] = ordered.empty() ? nullptr
                                 : reinterpret_cast<int*>(arena_.upload(ordered.data(), ordered.size(), stream_));
  }
}

// Point the single-step kernels at E-side psi level lvl.
template <typename real>
void Stepper<real>::set_psi_e(int lvl) {
  psi_e_cur_ = lvl;
  real* base = psi_e_lvl_[lvl];
  for (int p = 0; p < 2; ++p) {
    ma_[p].psiE = base;
    for (int s = 0; s < 6; ++s)
      fa_[p].psiE[s] = args_.psi[s] ? base + (args_.psi[s] - psi_side_[0]) : nullptr;
  }
}

// Steps step and step+1 in one launch pair (psi-free tiles, CPML tiles).
// Fields and H-side psi go from level cur_ to the other level, E-side psi
// and the ADE polarisation between their own two levels. DFT samples of
// step are written by the kernel into the ring slot reserved here; step+1
// is captured from the output level as usual.
template <typename real>
void Stepper<real>::step_pair(int64_t step) {
  const int in = cur_, out = cur_ ^ 1;
  pfhip::MarchArgs2<real> a = m2_;
  a.Lin = lvl_[in];
  a.Lout = lvl_[out];
  a.psiHin = in == 0 ? psi_side_[1] : psi_h_lvl1_;
  a.psiHout = in == 0 ? psi_h_lvl1_ : psi_side_[1];
  a.psiEin = psi_e_lvl_[psi_e_cur_];
  a.psiEout = psi_e_lvl_[psi_e_cur_ ^ 1];
  a.step = step;
  a.ade = ade_dev_[p_lvl_];
  for (const Dft& d : dft_) {
    if (step % d.interval != 0) continue;
    const int64_t sample = step / d.interval;
    if (sample >= d.n_samples) throw ArgError{"DFT sample beyond the phase table"};
    if (d.count != 0 && sample != d.sample0 + d.count) throw ArgError{"DFT samples must be consecutive"};
  }
  a.dft_slots = 0xFFFFFFFFu;
  for (size_t m = 0; m < dft_.size(); ++m) {
    Dft& d = dft_[m];
    if (step % d.interval != 0) continue;
    if (d.count == 0) d.sample0 = step / d.interval;
    a.dft_slots &= ~(0xFFu << (8 * m));
    a.dft_slots |= (uint32_t)d.count << (8 * m);
    ++d.count;
  }
  timed(PF_K_YEE_PAIR, [&] {
    // CPML tiles on the side stream when both lists are in use (disjoint
    // outputs, shared read-only inputs), overlapping the psi-free launch
    const bool fork = side_ && ntiles2_[0] > 0 && ntiles2_[1] > 0;
    if (fork) {
      HIP_CHECK(hipEventRecord(fork_ev_[0], stream_));
      HIP_CHECK(hipStreamWaitEvent(side_, fork_ev_[0], 0));
    }
    for (int q = 1; q >= 0; --q) {
      a.tiles = tiles2_[q];
      a.ntiles = ntiles2_[q];
      pfhip::launch_march2<real>(a, exact_, q == 1, pair_nw_, fork && q == 1 ? side_ : stream_);
    }
    if (fork) {
      HIP_CHECK(hipEventRecord(fork_ev_[1], side_));
      HIP_CHECK(hipStreamWaitEvent(stream_, fork_ev_[1], 0));
    }
  });
  for (Ade& m : ade_) {
    std::swap(m.Pcur, m.Pcur1);
    std::swap(m.Pprev, m.Pprev1);
  }
  p_lvl_ ^= 1;
  set_psi_e(psi_e_cur_ ^ 1);
  cur_ = out;
  timed(PF_K_AUX, [&] {
    for (Dft& d : dft_)
      if (d.count == d.K) fold(d);
    capture_dft(step + 1);
  });
}

template <typename real>
hipEvent_t Stepper<real>::record_event() {
  hipEvent_t e;
  if (pool_.empty()) {
    HIP_CHECK(hipEventCreate(&e));
  } else {
    e = pool_.back();
    pool_.pop_back();
  }
  HIP_CHECK(hipEventRecord(e, stream_));
  return e;
}

// Brackets the launches of `launch` with events on the handle's stream and
// books the interval to kernel class `cls`. A group opened inside another
// (a DFT fold issued by a capture in the aux group) closes the enclosing
// group first and reopens it afterwards, so no interval is counted twice.
template <typename real>
template <typename F>
void Stepper<real>::timed(int cls, F&& launch) {
  if (!timing_ || (skip_step_groups_ && (cls == PF_K_YEE_FUSED || cls == PF_K_AUX) && open_cls_ < 0)) {
    launch();
    if (timing_) chain_ev_ = nullptr;   // untimed launches follow any nested group's end
    return;
  }
  const int outer = open_cls_;
  if (outer >= 0) pending_.push_back({open_ev_, record_event(), outer, false});
  open_cls_ = cls;
  const bool chained = outer < 0 && chain_ev_;
  open_ev_ = chained ? chain_ev_ : record_event();
  chain_ev_ = nullptr;
  launch();
  hipEvent_t close = record_event();
  pending_.push_back({open_ev_, close, cls, chained});
  open_cls_ = outer;
  if (outer >= 0) open_ev_ = record_event();
  else chain_ev_ = close;
}

template <typename real>
void Stepper<real>::collect_timing() {
  chain_ev_ = nullptr;
  if (pending_.empty()) return;
  HIP_CHECK(hipStreamSynchronize(stream_));
  for (Timed& t : pending_) {
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, t.a, t.b));
    t_ms_[t.cls] += ms;
    t_count_[t.cls] += 1;
    if (!t.shared_start) pool_.push_back(t.a);   // a shared start is the previous entry's end
    pool_.push_back(t.b);
  }
  pending_.clear();
}

template <typename real>
void Stepper<real>::timing_enable(int every) {
  collect_timing();
  timing_ = every > 0;
  timing_every_ = std::max(every, 1);
  skip_step_groups_ = false;
  for (int c = 0; c < PF_NKCLASS; ++c) {
    t_ms_[c] = 0.0;
    t_count_[c] = 0;
  }
}

template <typename real>
void Stepper<real>::timing_read(double* ms, int64_t* launches) {
  collect_timing();
  for (int c = 0; c < PF_NKCLASS; ++c) {
    if (ms) ms[c] = t_ms_[c];
    if (launches) launches[c] = t_count_[c];
  }
}

// One step of the march over every tile list, reading and writing the
// levels `base` names (the stepper's ping-pong levels, or trajectory slots of
// the adjoint's recompute).
template <typename real>
void Stepper<real>::launch_march_lists(const pfhip::MarchArgs<real>& base) {
  // The CPML lists in side_mask_ on the side stream when the main
  // stream has tiles too, so the launches overlap (they write disjoint
  // tiles of level out and only read level in); the slowest list goes
  // first. With the mixed launch, the main stream's other lists follow it.
  static constexpr int order[NLIST] = {1, 4, 3, 2, 0};
  auto on_side = [&](int q) { return (side_mask_ >> q & 1) != 0; };
  bool any_side = false, any_main = false;
  for (int q = 0; q < NLIST; ++q)
    if (!(q != 0 && (mix_mask_ >> q & 1))) (on_side(q) ? any_side : any_main) |= ntiles_[q] > 0;
  const bool fork = side_ && any_side && any_main;
  if (fork) {
    HIP_CHECK(hipEventRecord(fork_ev_[0], stream_));
    HIP_CHECK(hipStreamWaitEvent(side_, fork_ev_[0], 0));
  }
  int seq[NLIST], nseq = 0;
  if (fork && mix_mask_) {
    for (int q : order)
      if (on_side(q)) seq[nseq++] = q;
    seq[nseq++] = 0;
    for (int q : order)
      if (q != 0 && !on_side(q)) seq[nseq++] = q;
  } else {
    for (int q : order) seq[nseq++] = q;
  }
  for (int n = 0; n < nseq; ++n) {
    const int q = seq[n];
    if (ntiles_[q] == 0 || (mix_mask_ >> q & 1)) continue;
    pfhip::MarchArgs<real> ma = base;
    ma.tiles = tiles_[q];
    ma.ntiles = ntiles_[q];
    hipStream_t s = fork && on_side(q) ? side_ : stream_;
    if (q == 0 && mix_mask_) {
      ma.tiles = tiles_mix_;
      ma.ntiles = ntiles_mix_;
      pfhip::launch_march64_mix<real>(ma, exact_, 12, (mix_mask_ >> 3 & 1) << 1 | (mix_mask_ >> 4 & 1), s);
    } else if (zpad_ == pfhip::ZPAD_BLOCK)
      pfhip::launch_march64<real>(ma, exact_, list_mask_[q], q == 0 ? march_nw_ : pml_nw_, s);
    else
      pfhip::launch_march<real>(ma, exact_, q == 1, q == 1 ? pml_nw_ : march_nw_,
                                march_dist_, s);
  }
  if (fork) {
    HIP_CHECK(hipEventRecord(fork_ev_[1], side_));
    HIP_CHECK(hipStreamWaitEvent(stream_, fork_ev_[1], 0));
  }
}

template <typename real>
void Stepper<real>::step_once(int64_t step, bool monitors) {
  if (fused_) {
    const int in = cur_;
    timed(PF_K_YEE_FUSED, [&] {
      if (fa_[in].variant == 2) {
        launch_march_lists(ma_[in]);
      } else {
        pfhip::launch_fused<real>(fa_[in], exact_, stream_);
      }
    });
    after_march(step);
    return;
  } else {
    timed(PF_K_YEE_H, [&] { pfhip::launch_generic_h<real>(args_, exact_, stream_); });
    timed(PF_K_AUX, [&] {
      pfhip::launch_inject<real>(fields_dev_[0], src_h_, n_steps_, step, stream_);
      for (Ade& a : ade_)
        pfhip::launch_ade_capture<real>(fields_dev_[0], a.cells, a.n, a.En, stream_);
    });
    timed(PF_K_YEE_E, [&] { pfhip::launch_generic_e<real>(args_, exact_, stream_); });
    timed(PF_K_AUX, [&] {
      for (Ade& a : ade_)
        pfhip::launch_ade_apply<real>(fields_dev_[0], a.cells, a.n, a.En, a.Pcur, a.Pprev, a.abc,
                                      a.n_poles, a.inv_eps, stream_);
      pfhip::launch_inject<real>(fields_dev_[0], src_e_, n_steps_, step, stream_);
      for (Flux& f : flux_)
        if (monitors)
          pfhip::launch_flux<real>(args_.f, f.axis, f.index, args_.nx, args_.ny, args_.nz, args_.nz, 0, f.dA,
                                   f.dt, f.partial, f.n_part, f.total, stream_);
    });
  }
  if (monitors) timed(PF_K_AUX, [&] { capture_dft(step); });
}

// Rest of a fused step once the march has written level cur_ ^ 1: ADE
// polarisation update, E-side sources, flux planes and DFT samples.
template <typename real>
void Stepper<real>::after_march(int64_t step) {
  const int in = cur_, out = cur_ ^ 1;
  pfhip::AdeJob<real>*& tab = ade_tab1_[in][p_lvl_];
  bool any_poles = false;
  for (const Ade& a : ade_) any_poles |= a.n > 0 && a.n_poles > 0;
  if ((ade_.size() > 1 || post_ok_) && any_poles && !tab) {
    // ade_jobs reads the polarisation pointers of the current P level
    std::vector<pfhip::AdeJob<real>> jobs;
    ade_jobs(in, jobs);
    ade_n_jobs1_ = (int)jobs.size();
    ade_job_n_.clear();
    for (const pfhip::AdeJob<real>& J : jobs) {
      ade_max_n1_ = std::max(ade_max_n1_, J.n);
      ade_job_n_.push_back(J.n);
    }
    tab = jobs.empty() ? nullptr : arena_.upload(jobs.data(), jobs.size(), stream_);
    chain_ev_ = nullptr;   // the upload is not part of the next group
  }
  if (post_ok_) {
    timed(PF_K_AUX, [&] { post_step(step, tab); });
    return;
  }
  timed(PF_K_AUX, [&] {
    if (tab && ade_n_jobs1_ > 0) {
      // every medium in one launch (the same per-node code as k_ade_pp)
      pfhip::launch_ade_jobs<real>(tab, ade_n_jobs1_, ade_max_n1_, stream_);
    } else {
      for (Ade& a : ade_)
        pfhip::launch_ade_pp<real>(f_[in], f_[out], a.cells, a.n, a.Pcur, a.Pprev, a.abc,
                                   a.n_poles, a.inv_eps, stream_);
    }
    cur_ = out;
    pfhip::launch_inject<real>(fields_dev_[cur_], src_e_, n_steps_, step, stream_);
    for (Flux& f : flux_)
      pfhip::launch_flux<real>(f_[cur_], f.axis, f.index, args_.nx, args_.ny, args_.nz, zrow_,
                               zpad_, f.dA, f.dt, f.partial, f.n_part, f.total, stream_);
    capture_dft(step);   // one timed group: two fewer events per step
  });
}

// after_march in one launch (k_post_step): ADE, E-side sources and the DFT
// samples of the step, then the flux planes and any full DFT ring's fold.
template <typename real>
void Stepper<real>::post_step(int64_t step, const pfhip::AdeJob<real>* tab) {
  const int out = cur_ ^ 1;
  cur_ = out;
  pfhip::PostArgs<real> P;
  std::memset(&P, 0, sizeof(P));
  auto blocks = [](int64_t n) { return (int)((n + 255) / 256); };
  P.ade = tab;
  P.n_ade = tab ? ade_n_jobs1_ : 0;
  for (int j = 0; j < P.n_ade; ++j) P.ade_blk[j + 1] = P.ade_blk[j] + blocks(ade_job_n_[j]);
  P.src_fields = fields_dev_[out];
  P.src = src_e_;
  P.n_steps = n_steps_;
  P.step = step;
  P.src_blk = blocks(src_e_.n_groups);
  P.dft_fields = fields_dev_[out];
  P.plane = (int64_t)args_.ny * zrow_;
  P.row = zrow_;
  P.zpad = zpad_;
  Dft* taken[pfhip::POST_MAX_DFT];
  for (Dft& m : dft_) {
    if (!dft_take(m, step)) continue;
    pfhip::PostGather<real>& G = P.dft[P.n_dft];
    G.comps = m.comps_dev;
    G.box = m.box;
    G.nbox = m.nbox;
    G.blocks_c = blocks(m.nbox);
    G.out = m.ring + (int64_t)m.count * m.n_comp * m.nbox;
    P.dft_blk[P.n_dft + 1] = P.dft_blk[P.n_dft] + m.n_comp * G.blocks_c;
    taken[P.n_dft++] = &m;
  }
  pfhip::launch_post_step<real>(P, stream_);
  for (Flux& f : flux_)
    pfhip::launch_flux<real>(f_[cur_], f.axis, f.index, args_.nx, args_.ny, args_.nz, zrow_,
                             zpad_, f.dA, f.dt, f.partial, f.n_part, f.total, stream_);
  for (int d = 0; d < P.n_dft; ++d)
    if (++taken[d]->count == taken[d]->K) fold(*taken[d]);
}

// A simulation can join a batched 3-D march when it runs the march kernel
// (fused ping-pong layout on CPML slabs, no H-side sources) and has no flux
// planes (a batch has no call to read them).
template <typename real>
void Stepper<real>::check_march_batch() const {
  if (shape[2] < 2) throw ArgError{"3-D batched stepping needs nz >= 2"};
  if (!fused_ || fa_[0].variant != 2)
    throw ArgError{"3-D batched stepping needs the march kernel: no H-side sources and CPML "
                   "slabs of the form [0, lo) U [hi, n)"};
  if (!flux_.empty()) throw ArgError{"flux monitors are not supported in a batch"};
  if (pair_) throw ArgError{"3-D batched stepping uses the single-step march (unset PHOTONFDTD_HIP_PAIR)"};
}

// March tiles of this grid for workgroups of nw waves and x-chunks of
// xchunk planes, split as the single run splits them: cpml false gives the
// psi-free tiles, true the tiles that touch a CPML slab (in x chunks of at
// most 16 planes). With dense rows every tile is a CPML tile.
template <typename real>
std::vector<pfhip::MarchTile> Stepper<real>::march_tile_list(int nw, int xchunk, bool cpml) const {
  const int n[3] = {args_.nx, args_.ny, args_.nz};
  const int len[3] = {xchunk, nw - 2, zpad_ ? zpad_ : 62};
  std::vector<Seg> seg[3];
  for (int a = 0; a < 3; ++a) {
    const int b0 = std::max(slab_e_[a][0], slab_h_[a][0] + 1);
    const int b1 = std::min(slab_e_[a][1], slab_h_[a][1]);
    if (a > 0 && zpad_ == pfhip::ZPAD_BLOCK) {
      seg[a] = split_aligned(n[a], b0, b1, a == 2 ? pfhip::ZPAD_BLOCK : len[a]);
    } else if (a == 2 && zpad_) {
      for (int z = 0; z < n[2]; z += zpad_) {
        const int c = std::min(zpad_, n[2] - z);
        seg[2].push_back({z, c, z >= b0 && z + c <= b1});
      }
    } else {
      seg[a] = zpad_ ? split_axis(n[a], b0, b1, len[a], a > 0) : split_axis(n[a], 0, 0, len[a], false);
    }
  }
  std::vector<pfhip::MarchTile> out;
  for (const Seg& sx : seg[0])
    for (const Seg& sy : seg[1])
      for (const Seg& sz : seg[2]) {
        const bool bulk = zpad_ && sx.bulk && sy.bulk && sz.bulk;
        if (bulk == cpml) continue;
        if (!cpml) {
          out.push_back({sx.start, sx.start + sx.count, sy.start, sy.count, sz.start, sz.count, 0, 0});
          continue;
        }
        const int px = zpad_ ? 16 : sx.count;
        for (int x = sx.start; x < sx.start + sx.count; x += px)
          out.push_back({x, std::min(sx.start + sx.count, x + px), sy.start, sy.count, sz.start,
                         sz.count, 0, 0});
      }
  check_march_tiles(out, nw);
  return out;
}

// The single-step march kernels assume these tile bounds (their lane and
// wave roles, the LDS plane table and the z block of each layout); a tile
// outside them is a host bug and never reaches the device.
template <typename real>
void Stepper<real>::check_march_tiles(const std::vector<pfhip::MarchTile>& tiles, int nw) const {
  const int nx = args_.nx, ny = args_.ny, nz = args_.nz;
  const int zmax = zpad_ ? zpad_ : 62;
  for (const pfhip::MarchTile& t : tiles) {
    const bool ok = t.i0 >= 0 && t.i1 > t.i0 && t.i1 <= nx && t.i1 - t.i0 <= pfhip::MARCH_MAX_XCHUNK &&
                    t.j0 >= 0 && t.jn >= 1 && t.jn <= nw - 2 && t.j0 + t.jn <= ny && t.k0 >= 0 &&
                    t.kn >= 1 && t.kn <= zmax && t.k0 + t.kn <= nz && (!zpad_ || t.k0 % zpad_ == 0);
    if (!ok) throw HipError{PF_E_STATE, "march tile outside the kernel's bounds"};
  }
}

template <typename real>
void Stepper<real>::ade_jobs(int p, std::vector<pfhip::AdeJob<real>>& out) const {
  for (const Ade& a : ade_) {
    if (a.n <= 0 || a.n_poles == 0) continue;
    pfhip::AdeJob<real> J;
    for (int c = 0; c < 3; ++c) {
      J.Ein[c] = f_[p][c];
      J.Eout[c] = f_[p ^ 1][c];
    }
    J.cells = a.cells;
    J.n = a.n;
    J.Pcur = a.Pcur;
    J.Pprev = a.Pprev;
    J.abc = a.abc;
    J.n_poles = a.n_poles;
    J.inv_eps = a.inv_eps;
    out.push_back(J);
  }
}

template <typename real>
bool Stepper<real>::src_job(int p, pfhip::SrcJob<real>& out) const {
  if (src_e_.n_groups == 0) return false;
  out.fields = fields_dev_[p ^ 1];
  out.g = src_e_;
  out.n_steps = n_steps_;
  return true;
}

template <typename real>
void Stepper<real>::finish_step(int64_t step) {
  cur_ ^= 1;
  timed(PF_K_AUX, [&] { capture_dft(step); });
}

template <typename real>
void Stepper<real>::dft_jobs(int p, int64_t step0, std::vector<pfhip::DftJob<real>>& out) const {
  for (const Dft& m : dft_) {
    pfhip::DftJob<real> J;
    J.fields = fields_dev_[p];
    J.comps = m.comps_dev;
    J.n_comp = m.n_comp;
    J.box = m.box;
    J.plane = (int64_t)args_.ny * zrow_;
    J.row = zrow_;
    J.zpad = zpad_;
    J.ring = m.ring;
    J.nbox = m.nbox;
    J.K = m.K;
    J.interval = m.interval;
    J.q0 = (step0 + m.interval - 1) / m.interval;
    J.count0 = m.count;
    out.push_back(J);
  }
}

// The ring bookkeeping of capture_dft for a sample the batched launch wrote.
template <typename real>
void Stepper<real>::finish_step_dft(int64_t step) {
  cur_ ^= 1;
  for (Dft& m : dft_) {
    if (step % m.interval != 0) continue;
    const int64_t sample = step / m.interval;
    if (sample >= m.n_samples) throw ArgError{"DFT sample beyond the phase table"};
    if (m.count == 0) m.sample0 = sample;
    else if (sample != m.sample0 + m.count) throw ArgError{"DFT samples must be consecutive"};
    if (++m.count == m.K) fold(m);
  }
}

// Run on the stream of a batch from now on (the batch owns and destroys it).
template <typename real>
void Stepper<real>::adopt_stream(hipStream_t s) {
  HIP_CHECK(hipStreamSynchronize(stream_));
  if (own_stream_) HIP_CHECK(hipStreamDestroy(stream_));
  stream_ = s;
  own_stream_ = false;
}

template <typename real>
void Stepper<real>::save_state(real* dst) {
  const size_t nf = 6 * (size_t)(shape[0] * shape[1] * shape[2]);
  // the fused stepper: its current level (dense rows of nz slots, checked
  // by checkpoint_enable) and that level's H-side psi
  const real* lv = fused_ ? lvl_[cur_] : lvl_[0];
  const real* ph = fused_ && cur_ == 1 ? psi_h_lvl1_ : psi_side_[1];
  HIP_CHECK(hipMemcpyAsync(dst, lv, nf * sizeof(real), hipMemcpyDeviceToDevice, stream_));
  if (psi_len_e_)
    HIP_CHECK(hipMemcpyAsync(dst + nf, psi_side_[0], psi_len_e_ * sizeof(real),
                             hipMemcpyDeviceToDevice, stream_));
  if (psi_len_h_)
    HIP_CHECK(hipMemcpyAsync(dst + nf + psi_len_e_, ph, psi_len_h_ * sizeof(real),
                             hipMemcpyDeviceToDevice, stream_));
}

template <typename real>
void Stepper<real>::describe(pfhip::SimDev<real>& s, int pol, PostNodes& post) {
  if (fused_) throw ArgError{"batched stepping needs the generic layout (PF_FLAG_GENERIC)"};
  if (shape[2] != 1 || shape[0] < 2 || shape[1] < 2)
    throw ArgError{"batched stepping supports xy-plane 2-D grids (nx, ny >= 2, nz == 1)"};
  if (shape[0] * shape[1] >= (int64_t(1) << 31)) throw ArgError{"grid too large for a batch"};
  if (pol != 0 && pol != 1) throw ArgError{"polarisation must be 0 (TE) or 1 (TM)"};
  if (!slab_ok_) throw ArgError{"batched stepping needs CPML slabs of the form [0, lo) U [hi, n)"};
  if (dft_.size() > (size_t)pfhip::BATCH_MAX_DFT) throw ArgError{"too many DFT monitors for a batch"};
  if (ade_.size() > (size_t)pfhip::BATCH_MAX_ADE) throw ArgError{"too many dispersive media for a batch"};
  if (!flux_.empty()) throw ArgError{"flux monitors are not supported in a batch"};
  const int ny = (int)shape[1];
  // Two time levels with padded y rows for the batch march (each 62-column
  // strip owns an aligned 64-slot block); this stepper's own generic
  // kernels never run on them.
  if (!ypad_) {
    ypad_ = pfhip::MARCH_OWNED;
    yrow_ = pfhip::zpad_row(ny, ypad_);
    fcell_ = shape[0] * yrow_;
    if (fcell_ >= (int64_t(1) << 31)) throw ArgError{"grid too large for a batch"};
    for (int p = 0; p < 2; ++p) {
      if (lvl_[p]) arena_.free(lvl_[p]);
      lvl_[p] = arena_.zeros<real>(6 * (size_t)fcell_, stream_);
      for (int c = 0; c < 6; ++c) f_[p][c] = lvl_[p] + (size_t)c * fcell_;
      fields_dev_[p] = arena_.upload(f_[p], 6, stream_);
    }
    for (int c = 0; c < 6; ++c) args_.f[c] = nullptr;
    std::vector<real> ce((size_t)ncell());
    copy_sync(ce.data(), args_.ce, ce.size() * sizeof(real), hipMemcpyDeviceToHost, stream_);
    const std::vector<real> cep = pad_rows(ce.data(), (int)shape[0], 1, ny, yrow_, ypad_);
    ce_f_ = arena_.upload(cep.data(), cep.size(), stream_);
  }
  // dense node index i * ny + j -> its slot in the padded rows
  auto slot = [&](int64_t c) { return (int)((c / ny) * yrow_ + pfhip::zslot(c % ny, ypad_)); };
  if (!psi_h_lvl1_ && psi_len_h_) psi_h_lvl1_ = arena_.zeros<real>(psi_len_h_, stream_);
  std::memset(&s, 0, sizeof(s));
  for (int c = 0; c < 6; ++c) {
    s.f[0][c] = f_[0][c];
    s.f[1][c] = f_[1][c];
  }
  for (int q = 0; q < 6; ++q) {
    s.psiE[q] = args_.psi[q];
    real* h0 = args_.psi[6 + q];
    s.psiH[0][q] = h0;
    s.psiH[1][q] = h0 ? psi_h_lvl1_ + (h0 - psi_side_[1]) : nullptr;
  }
  s.ce = ce_f_;
  s.nyr = yrow_;
  for (int a = 0; a < 2; ++a) {
    s.b_e[a] = args_.b_e[a];
    s.c_e[a] = args_.c_e[a];
    s.b_h[a] = args_.b_h[a];
    s.c_h[a] = args_.c_h[a];
    for (int q = 0; q < 2; ++q) {
      s.slab_e[a][q] = slab_e_[a][q];
      s.slab_h[a][q] = slab_h_[a][q];
    }
    s.pe[a] = args_.pe[a];
    s.ph[a] = args_.ph[a];
    s.d[a] = args_.d[a];
    s.inv_d[a] = args_.inv_d[a];
  }
  s.nx = args_.nx;
  s.ny = args_.ny;
  s.ch = args_.ch;
  s.pol = pol;
  s.n_steps = n_steps_;
  s.src_e = src_e_;
  s.src_h = src_h_;
  s.src_lo = INT64_MAX;
  s.src_hi = -1;
  for (int side = 0; side < 2; ++side) {
    const pfhip::SourceGroupsDev& g = side ? src_h_ : src_e_;
    if (g.n_groups == 0) continue;
    std::vector<int64_t> cells(g.n_groups);
    std::vector<int32_t> comps(g.n_groups);
    copy_sync(cells.data(), g.cell, cells.size() * sizeof(int64_t), hipMemcpyDeviceToHost, stream_);
    copy_sync(comps.data(), g.comp, comps.size() * sizeof(int32_t), hipMemcpyDeviceToHost, stream_);
    for (int q = 0; q < g.n_groups; ++q) {
      const bool active = pol == 0 ? (comps[q] == 0 || comps[q] == 1 || comps[q] == 5)
                                   : (comps[q] == 2 || comps[q] == 3 || comps[q] == 4);
      if (!active) throw ArgError{"a source excites the idle polarisation group"};
      if (side == 0) post.esrc.push_back({slot(cells[q]), q});
    }
    if (side == 1) {
      s.src_lo = *std::min_element(cells.begin(), cells.end());
      s.src_hi = *std::max_element(cells.begin(), cells.end());
    }
  }
  for (size_t m = 0; m < ade_.size(); ++m) {
    if (ade_[m].n >= (int64_t(1) << 26)) throw ArgError{"dispersive region too large for a batch"};
    std::vector<int64_t> cells(ade_[m].n);
    copy_sync(cells.data(), ade_[m].cells, cells.size() * sizeof(int64_t), hipMemcpyDeviceToHost, stream_);
    for (int64_t t = 0; t < ade_[m].n; ++t)
      post.ade.push_back({slot(cells[t]), (int)(m << 26) | (int)t});
    pfhip::AdeDev<real>& A = s.ade[m];
    A.n = ade_[m].n;
    A.n_poles = ade_[m].n_poles;
    A.abc = ade_[m].abc;
    A.inv_eps = ade_[m].inv_eps;
    A.Pcur = ade_[m].Pcur;
    A.Pprev = ade_[m].Pprev;
  }
  for (size_t m = 0; m < dft_.size(); ++m) {
    const Dft& M = dft_[m];
    pfhip::DftDev& d = s.dft[m];
    d.n_comp = M.n_comp;
    for (int c = 0; c < 6; ++c) d.comps[c] = c < M.n_comp ? M.comps[c] : -1;
    d.nbox = M.nbox;
    d.n_freq = M.n_freq;
    d.interval = M.interval;
    d.phase_e = M.phase_e;
    d.phase_h = M.phase_h;
    d.dt = M.dt;
    d.acc = M.acc;
    for (int c = 0; c < M.n_comp; ++c) {
      const int comp = M.comps[c];
      const bool active = pol == 0 ? (comp == 0 || comp == 1 || comp == 5)
                                   : (comp == 2 || comp == 3 || comp == 4);
      if (!active) continue;   // idle components stay zero: their accumulators too
      for (int64_t bi = 0; bi < M.box.count[0]; ++bi)
        for (int64_t bj = 0; bj < M.box.count[1]; ++bj) {
          const int i = (int)(M.box.start[0] + bi * M.box.stride[0]);
          const int j = (int)(M.box.start[1] + bj * M.box.stride[1]);
          post.dft.push_back({slot((int64_t)i * ny + j), (int)m, c, comp, bi * M.box.count[1] + bj});
        }
    }
  }
  HIP_CHECK(hipStreamSynchronize(stream_));
}

template <typename real>
void Stepper<real>::reset(hipStream_t s) {
  const size_t nf = 6 * (size_t)fcell_;
  for (int p = 0; p < 2; ++p)
    if (lvl_[p]) HIP_CHECK(hipMemsetAsync(lvl_[p], 0, nf * sizeof(real), s));
  if (psi_len_e_) HIP_CHECK(hipMemsetAsync(psi_side_[0], 0, psi_len_e_ * sizeof(real), s));
  if (psi_len_h_) HIP_CHECK(hipMemsetAsync(psi_side_[1], 0, psi_len_h_ * sizeof(real), s));
  if (psi_h_lvl1_) HIP_CHECK(hipMemsetAsync(psi_h_lvl1_, 0, psi_len_h_ * sizeof(real), s));
  if (psi_e_lvl_[1]) HIP_CHECK(hipMemsetAsync(psi_e_lvl_[1], 0, psi_len_e_ * sizeof(real), s));
  if (pair_) set_psi_e(0);
  if (p_lvl_) {
    for (Ade& a : ade_) {
      std::swap(a.Pcur, a.Pcur1);
      std::swap(a.Pprev, a.Pprev1);
    }
    p_lvl_ = 0;
  }
  for (Ade& a : ade_) {
    const size_t n = 3 * (size_t)a.n_poles * a.n;
    if (n) {
      for (real* P : {a.Pcur, a.Pprev, a.Pcur1, a.Pprev1})
        if (P) HIP_CHECK(hipMemsetAsync(P, 0, n * sizeof(real), s));
    }
  }
  for (Dft& m : dft_) {
    HIP_CHECK(hipMemsetAsync(m.acc, 0, (size_t)m.n_comp * m.n_freq * m.nbox * 2 * sizeof(double), s));
    m.count = 0;
    m.sample0 = 0;
  }
  cur_ = 0;
}

template <typename real>
void Stepper<real>::checkpoint_enable(int interval) {
  // the fused stepper qualifies when its fields are in the generic layout
  // (dense z rows of exactly nz slots) and it steps one step at a time
  if (fused_ && !(fa_[0].variant == 2 && zpad_ == pfhip::ZPAD_BLOCK && zrow_ == shape[2] && !pair_))
    throw ArgError{"checkpoints need the generic layout: create with PF_FLAG_GENERIC, or with "
                   "PF_FLAG_DENSE64 on a grid whose nz is a multiple of 64"};
  if (args_.ce_y) throw ArgError{"the adjoint does not support subpixel smoothing"};
  if (interval < 1) throw ArgError{"checkpoint interval must be positive"};
  if (!ade_.empty()) throw ArgError{"the adjoint does not support dispersive media"};
  ckpt_interval_ = interval;
  const size_t n = (size_t)((n_steps_ + interval - 1) / interval);
  const size_t elems = 6 * (size_t)(shape[0] * shape[1] * shape[2]) + psi_len_e_ + psi_len_h_;
  ckpts_.clear();
  for (size_t c = 0; c < n; ++c) ckpts_.push_back(static_cast<real*>(arena_.alloc(elems * sizeof(real))));
}

template <typename real>
void Stepper<real>::adjoint(int n_cot, const pf_cotangent* cot, double* grad_out) {
  if (ckpt_interval_ <= 0) throw ArgError{"pf_checkpoint_enable was not called before the forward run"};
  if (steps_done_ != n_steps_) throw ArgError{"run the forward pass to the last step first"};
  if (!grad_out) throw ArgError{"null gradient destination"};
  const int64_t ncell = shape[0] * shape[1] * shape[2];
  const size_t nf = 6 * (size_t)ncell;
  const size_t elems = nf + psi_len_e_ + psi_len_h_;
  const int C = ckpt_interval_;
  DeviceArena scratch;   // freed when the sweep ends
  // cotangents: per monitor [n_comp][n_freq][nbox][2], zero where not given
  std::vector<double*> cot_dev(dft_.size(), nullptr);
  for (size_t m = 0; m < dft_.size(); ++m)
    cot_dev[m] = scratch.zeros<double>((size_t)dft_[m].n_comp * dft_[m].n_freq * dft_[m].nbox * 2, stream_);
  for (int c = 0; c < n_cot; ++c) {
    const pf_cotangent& g = cot[c];
    if (g.mon < 0 || g.mon >= (int)dft_.size() || g.slot < 0 || g.slot >= dft_[g.mon].n_comp || !g.g)
      throw ArgError{"bad cotangent"};
    const Dft& m = dft_[g.mon];
    const size_t n = (size_t)m.n_freq * m.nbox * 2;
    copy_sync(cot_dev[g.mon] + (size_t)g.slot * n, g.g, n * sizeof(double), hipMemcpyDefault,
              stream_);
  }
  real* lam = scratch.zeros<real>(elems, stream_);
  real* lam_d = scratch.zeros<real>(6 * (size_t)ncell, stream_);
  real* lam_dh = scratch.zeros<real>(6 * (size_t)ncell, stream_);
  double* grad = scratch.zeros<double>(ncell, stream_);
  // C + 1 trajectory slots: the states before every step of a segment and
  // after its last one
  real* traj = static_cast<real*>(scratch.alloc((size_t)(C + 1) * elems * sizeof(real)));
  real* lam_tab[6];
  for (int c = 0; c < 6; ++c) lam_tab[c] = lam + (size_t)c * ncell;
  real** lam_dev = scratch.upload(lam_tab, 6, stream_);

  pfhip::AdjArgs<real> a;
  std::memset(&a, 0, sizeof(a));
  for (int c = 0; c < 6; ++c) {
    a.f[c] = f_[0][c];
    a.lam_f[c] = lam_tab[c];
    a.lam_d[c] = lam_d + (size_t)c * ncell;
    a.lam_dh[c] = lam_dh + (size_t)c * ncell;
  }
  a.ce = args_.ce;
  for (int sl = 0; sl < 12; ++sl) {
    const size_t off = sl < 6 ? (size_t)(args_.psi[sl] - args_.psi[0])
                              : psi_len_e_ + (size_t)(args_.psi[sl] - args_.psi[6]);
    a.lam_psi[sl] = lam + nf + off;
    a.psi_fwd[sl] = args_.psi[sl];   // both sides reset per reverse step (trajectory)
  }
  a.grad = grad;
  for (int ax = 0; ax < 3; ++ax) {
    a.b_e[ax] = args_.b_e[ax];
    a.c_e[ax] = args_.c_e[ax];
    a.b_h[ax] = args_.b_h[ax];
    a.c_h[ax] = args_.c_h[ax];
    a.map_e[ax] = args_.map_e[ax];
    a.map_h[ax] = args_.map_h[ax];
    a.pe[ax] = args_.pe[ax];
    a.ph[ax] = args_.ph[ax];
    a.inv_d[ax] = args_.inv_d[ax];
  }
  a.nx = args_.nx;
  a.ny = args_.ny;
  a.nz = args_.nz;
  a.nyz = args_.nyz;
  a.ch = args_.ch;

  // field tables of the trajectory slots (the source kernels take one)
  std::vector<real*> slot_ptrs((size_t)(C + 1) * 6);
  for (int k = 0; k <= C; ++k)
    for (int c = 0; c < 6; ++c) slot_ptrs[(size_t)k * 6 + c] = traj + (size_t)k * elems + (size_t)c * ncell;
  real** slot_tab = scratch.upload(slot_ptrs.data(), slot_ptrs.size(), stream_);
  // psi slot offsets inside a stored state (E side, then H side)
  size_t psi_off[pfhip::PSI_COUNT];
  for (int sl = 0; sl < pfhip::PSI_COUNT; ++sl)
    psi_off[sl] = sl < 6 ? (size_t)(args_.psi[sl] - args_.psi[0])
                         : psi_len_e_ + (size_t)(args_.psi[sl] - args_.psi[6]);

  // The recompute runs the forward's own march when the stepper marches on
  // dense rows of exactly nz slots (the trajectory's layout) with no H-side
  // sources and no dispersive media: the slots are its levels, the E-side psi
  // (updated in place by the march) copied into the output slot first.
  const bool march_rc = fused_ && fa_[0].variant == 2 && zpad_ == pfhip::ZPAD_BLOCK &&
                        zrow_ == shape[2] && !pair_ && src_h_.n_groups == 0 && ade_.empty() &&
                        !(std::getenv("PHOTONFDTD_HIP_ADJ_MARCH") &&
                          std::getenv("PHOTONFDTD_HIP_ADJ_MARCH")[0] == '0');
  const int64_t nseg = (n_steps_ + C - 1) / C;
  bool first_reverse = true;   // no H-update terms to gather before the last step's transpose
  for (int64_t sg = nseg - 1; sg >= 0; --sg) {
    const int64_t s0 = sg * C, s1 = std::min<int64_t>(s0 + C, n_steps_);
    // the segment's states: slot 0 is the checkpoint; each step then runs
    // out of place from slot k into slot k + 1, the last one included (its
    // H^{n+1/2} is what the reverse step of that step reads)
    HIP_CHECK(hipMemcpyAsync(traj, ckpts_[sg], elems * sizeof(real), hipMemcpyDeviceToDevice, stream_));
    for (int64_t st = s0; st < s1; ++st) {
      const int k = (int)(st - s0);
      real* Tin = traj + (size_t)k * elems;
      real* Tout = Tin + elems;
      pfhip::StepArgs<real> sa = args_;
      for (int c = 0; c < 6; ++c) {
        sa.f[c] = Tin + (size_t)c * ncell;
        sa.fo[c] = Tout + (size_t)c * ncell;
      }
      for (int sl = 0; sl < pfhip::PSI_COUNT; ++sl) {
        sa.psi[sl] = args_.psi[sl] ? Tin + nf + psi_off[sl] : nullptr;
        sa.psio[sl] = args_.psi[sl] ? Tout + nf + psi_off[sl] : nullptr;
      }
      real* const* out_tab = slot_tab + (size_t)(k + 1) * 6;
      if (march_rc) {
        pfhip::MarchArgs<real> m = ma_[0];
        m.Lin = Tin;
        m.Lout = Tout;
        m.psiE = psi_len_e_ ? Tout + nf : nullptr;
        m.psiHin = psi_len_h_ ? Tin + nf + psi_len_e_ : nullptr;
        m.psiHout = psi_len_h_ ? Tout + nf + psi_len_e_ : nullptr;
        if (psi_len_e_)
          HIP_CHECK(hipMemcpyAsync(Tout + nf, Tin + nf, psi_len_e_ * sizeof(real),
                                   hipMemcpyDeviceToDevice, stream_));
        launch_march_lists(m);
        pfhip::launch_inject<real>(out_tab, src_e_, n_steps_, st, stream_);
        continue;
      }
      // the generic step's order: H, H sources, E, E sources (no dispersive media here)
      pfhip::launch_generic_h<real>(sa, exact_, stream_);
      pfhip::launch_inject<real>(out_tab, src_h_, n_steps_, st, stream_);
      pfhip::launch_generic_e<real>(sa, exact_, stream_);
      pfhip::launch_inject<real>(out_tab, src_e_, n_steps_, st, stream_);
    }
    for (int64_t st = s1 - 1; st >= s0; --st) {
      const real* T = traj + (size_t)(st - s0) * elems;
      // adjoint sources of the samples taken after this step
      for (size_t m = 0; m < dft_.size(); ++m) {
        const Dft& d = dft_[m];
        if (st % d.interval != 0) continue;
        pfhip::launch_adj_dft_source<real>(lam_dev, d.comps_dev, d.is_h_dev, d.n_comp, d.box,
                                           args_.nyz, args_.nz, cot_dev[m], d.n_freq, d.phase_e,
                                           d.phase_h, st / d.interval, d.dt, stream_);
      }
      // E^n from the pre-step slot, H^{n+1/2} (after the H sources) from the
      // next slot, where the recompute wrote it; both psi sides from the
      // pre-step slot
      for (int c = 0; c < 3; ++c) a.f[c] = T + (size_t)c * ncell;
      for (int c = 3; c < 6; ++c) a.f[c] = T + elems + (size_t)c * ncell;
      for (int sl = 0; sl < 6; ++sl) {
        a.psi_fwd[sl] = T + nf + (size_t)(args_.psi[sl] - args_.psi[0]);
        a.psi_fwd[6 + sl] = T + nf + psi_len_e_ + (size_t)(args_.psi[6 + sl] - args_.psi[6]);
      }
      // E transpose (after the H-update terms of step st + 1), then the H
      // transpose (after the E-update terms of step st); the H-update terms
      // of the first step are never gathered, since nothing reads lam_E after
      pfhip::launch_adjoint_step<real>(a, stream_, 0, !first_reverse);
      pfhip::launch_adjoint_step<real>(a, stream_, 1, true);
      first_reverse = false;
    }
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(grad_out, grad, (size_t)ncell * sizeof(double), hipMemcpyDefault, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  scratch.release();
}

template <typename real>
bool Stepper<real>::dft_take(Dft& m, int64_t step) {
  if (step % m.interval != 0) return false;
  const int64_t sample = step / m.interval;
  if (sample >= m.n_samples) throw ArgError{"DFT sample beyond the phase table"};
  if (m.count == 0) m.sample0 = sample;
  else if (sample != m.sample0 + m.count) throw ArgError{"DFT samples must be consecutive"};
  return true;
}

template <typename real>
void Stepper<real>::capture_dft(int64_t step) {
  for (Dft& m : dft_) {
    if (!dft_take(m, step)) continue;
    pfhip::launch_gather_box<real>(fields_dev_[cur_], m.comps_dev, m.n_comp, m.box,
                                   (int64_t)args_.ny * zrow_, zrow_, zpad_,
                                   m.ring + (int64_t)m.count * m.n_comp * m.nbox, stream_);
    if (++m.count == m.K) fold(m);
  }
}

template <typename real>
void Stepper<real>::fold(Dft& m) {
  if (m.count == 0) return;
  timed(PF_K_DFT_FOLD, [&] {
    pfhip::launch_dft_fold<real>(m.ring, m.nbox, m.n_comp, m.is_h_dev, m.count, m.sample0,
                                 m.phase_e, m.phase_h, m.n_freq, m.dt, m.acc, stream_);
  });
  m.sample0 += m.count;
  m.count = 0;
}

template <typename real>
void Stepper<real>::info(pf_info* out) const {
  std::memset(out, 0, sizeof(*out));
  out->path = !fused_ ? 0 : fa_[0].variant != 2 ? 1 : pair_ ? 3 : 2;
  out->zpad = zpad_;
  out->zrow = zrow_;
  out->march_nw = march_nw_;
  out->pml_nw = pml_nw_;
  out->tiles_bulk = ntiles_[0];
  out->cells_bulk = tile_cells_[0];
  for (int q = 1; q < NLIST; ++q) {
    out->tiles_cpml += ntiles_[q];
    out->cells_cpml += tile_cells_[q];
    if (ntiles_[q] > 0) out->cpml_lists |= 1 << (q - 1);
  }
  out->pair_tiles_bulk = ntiles2_[0];
  out->pair_tiles_cpml = ntiles2_[1];
  for (const Dft& m : dft_) out->dft_ring = std::max<int64_t>(out->dft_ring, m.K);
  out->post_fused = fused_ && post_ok_;
}

// Folds every buffered DFT sample into the complex128 accumulators and
// waits, so the accumulators hold every step advanced so far.
template <typename real>
void Stepper<real>::dft_flush() {
  for (Dft& m : dft_) fold(m);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(stream_));
}

template <typename real>
void Stepper<real>::advance(int64_t step0, int64_t n) {
  if (step0 < 0 || n < 0) throw ArgError{"negative step range"};
  if ((src_h_.n_groups || src_e_.n_groups) && step0 + n > n_steps_)
    throw ArgError{"step range exceeds the source waveform table"};
  for (int64_t s = step0; s < step0 + n;) {
    if (ckpt_interval_ > 0 && s % ckpt_interval_ == 0) {
      if (s / ckpt_interval_ >= (int64_t)ckpts_.size()) throw ArgError{"step beyond the checkpoint table"};
      save_state(ckpts_[s / ckpt_interval_]);
      chain_ev_ = nullptr;   // the copies are not part of the next group
    }
    if (pair_ && ckpt_interval_ == 0 && s + 2 <= step0 + n) {
      step_pair(s);
      s += 2;
    } else {
      skip_step_groups_ = timing_ && s % timing_every_ != 0;
      step_once(s);
      skip_step_groups_ = false;
      s += 1;
    }
    if (timing_ && pending_.size() > 4096) collect_timing();
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(stream_));
  collect_timing();
  steps_done_ = step0 + n;
}

template <typename real>
real* Stepper<real>::staging(size_t elems) {
  if (elems > stage_elems_) {
    stage_ = static_cast<real*>(arena_.alloc(elems * sizeof(real)));
    stage_elems_ = elems;
  }
  return stage_;
}

// Gathers one component over a strided box into the staging buffer (on the
// handle's stream, not waited for); *nbox receives the box's node count.
template <typename real>
const real* Stepper<real>::gather(int comp, const pf_box& box, int64_t* nbox) {
  if (comp < 0 || comp > 5) throw ArgError{"component out of range"};
  const BoxDev b = check_box(box, shape);
  *nbox = b.count[0] * b.count[1] * b.count[2];
  real* out = staging(*nbox);
  if (ypad_) {
    // padded y rows of a 2-D batch member: gather (x, 1, y) with y as the padded axis
    BoxDev t = b;
    t.start[1] = 0;
    t.count[1] = 1;
    t.stride[1] = 1;
    t.start[2] = b.start[1];
    t.count[2] = b.count[1];
    t.stride[2] = b.stride[1];
    pfhip::launch_gather_box<real>(fields_dev_[cur_], comp_index_dev_ + comp, 1, t, yrow_, yrow_, ypad_,
                                   out, stream_);
  } else {
    pfhip::launch_gather_box<real>(fields_dev_[cur_], comp_index_dev_ + comp, 1, b,
                                   (int64_t)args_.ny * zrow_, zrow_, zpad_, out, stream_);
  }
  HIP_CHECK(hipGetLastError());
  return out;
}

template <typename real>
void Stepper<real>::read_box(int comp, const pf_box& box, void* dst) {
  if (!dst) throw ArgError{"null destination"};
  int64_t nbox = 0;
  const real* out = gather(comp, box, &nbox);
  HIP_CHECK(hipMemcpyAsync(dst, out, nbox * sizeof(real), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
}

// A snapshot quantised on the device for compressed storage
// (storage.py:128-141): bits 8 -> int8, bits 16 -> float16, one per node.
template <typename real>
void Stepper<real>::read_box_q(int comp, const pf_box& box, int bits, void* dst, double* scale) {
  if (!dst || !scale) throw ArgError{"null destination"};
  if (bits != 8 && bits != 16) throw ArgError{"bits must be 8 or 16"};
  int64_t nbox = 0;
  const real* x = gather(comp, box, &nbox);
  constexpr int kMaxPart = 1024;
  if (!qpart_) qpart_ = static_cast<double*>(arena_.alloc((kMaxPart + 1) * sizeof(double)));
  const size_t bytes = (size_t)nbox * (bits / 8);
  if (bytes > qbuf_bytes_) {
    qbuf_ = arena_.alloc(bytes);
    qbuf_bytes_ = bytes;
  }
  pfhip::launch_quantize<real>(x, nbox, bits, qpart_, kMaxPart, qbuf_, qpart_ + kMaxPart, stream_);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(dst, qbuf_, bytes, hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipMemcpyAsync(scale, qpart_ + kMaxPart, sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
}

template <typename real>
void Stepper<real>::write_field(int comp, const void* src) {
  if (comp < 0 || comp > 5) throw ArgError{"component out of range"};
  if (!src) throw ArgError{"null source"};
  const int64_t ncell = shape[0] * shape[1] * shape[2];
  if (zpad_ || ypad_) {
    const std::vector<real> p =
        zpad_ ? pad_rows(static_cast<const real*>(src), (int)shape[0], (int)shape[1], (int)shape[2],
                         zrow_, zpad_)
              : pad_rows(static_cast<const real*>(src), (int)shape[0], 1, (int)shape[1], yrow_, ypad_);
    copy_sync(f_[cur_][comp], p.data(), p.size() * sizeof(real), hipMemcpyHostToDevice, stream_);
    return;
  }
  HIP_CHECK(hipMemcpyAsync(f_[cur_][comp], src, ncell * sizeof(real), hipMemcpyHostToDevice, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
}

template <typename real>
void Stepper<real>::dft_read(int mon, int slot, double* dst) {
  if (mon < 0 || mon >= (int)dft_.size()) throw ArgError{"DFT monitor index out of range"};
  Dft& m = dft_[mon];
  if (slot < 0 || slot >= m.n_comp) throw ArgError{"DFT component slot out of range"};
  if (!dst) throw ArgError{"null destination"};
  fold(m);
  const size_t n = (size_t)m.n_freq * m.nbox * 2;
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(dst, m.acc + (size_t)slot * n, n * sizeof(double),
                           hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
}

template <typename real>
void Stepper<real>::field_view(int comp, pf_view* out) {
  if (comp < 0 || comp > 5) throw ArgError{"field component out of range"};
  if (zpad_ && zpad_ != pfhip::ZPAD_BLOCK)
    throw HipError{PF_E_STATE, "fields are kept in padded 64-slot blocks; use pf_read_field"};
  HIP_CHECK(hipStreamSynchronize(stream_));
  std::memset(out, 0, sizeof(*out));
  out->ptr = f_[fused_ ? cur_ : 0][comp];
  out->dtype = sizeof(real) == 4 ? PF_F32 : PF_F64;
  out->ndim = 3;
  for (int a = 0; a < 3; ++a) out->shape[a] = shape[a];
  out->strides[0] = shape[1] * zrow_;
  out->strides[1] = zrow_;
  out->strides[2] = 1;
  out->stream = (uint64_t)(uintptr_t)stream_;
}

template <typename real>
void Stepper<real>::dft_view(int mon, pf_view* out) {
  if (mon < 0 || mon >= (int)dft_.size()) throw ArgError{"DFT monitor index out of range"};
  Dft& m = dft_[mon];
  fold(m);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(stream_));
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
  out->stream = (uint64_t)(uintptr_t)stream_;
}

template <typename real>
void Stepper<real>::flux_read(double* dst) {
  if (flux_.empty()) return;
  if (!dst) throw ArgError{"null destination"};
  for (size_t m = 0; m < flux_.size(); ++m)
    HIP_CHECK(hipMemcpyAsync(dst + m, flux_[m].total, sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
}

// ---------------------------------------------------------------------------
// Batch of 2-D simulations advanced together.
// ---------------------------------------------------------------------------
struct BatchBase {
  virtual ~BatchBase() = default;
  virtual void configure(int mode, int waves_per_block, int rows) = 0;
  virtual void reset() = 0;
  virtual void persistent_grid(int waves_per_block, int* per_cu, int* n_cu, int* grid) = 0;
  virtual void advance(int64_t n) = 0;
  virtual pf_handle* sim(int i) = 0;
  virtual int size() const = 0;
  virtual double last_ms() const = 0;
  virtual void write_ce(int i, const void* ce) = 0;
};

template <typename real>
class Batch final : public BatchBase {
 public:
  Batch(const pf_desc* descs, int n, const int* pols) : exact_(descs[0].exact != 0) {
    if (n < 1) throw ArgError{"empty batch"};
    for (int i = 1; i < n; ++i)
      if ((descs[i].exact != 0) != exact_) throw ArgError{"a batch needs one arithmetic mode"};
    HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    std::vector<pfhip::SimDev<real>> host(n);
    std::vector<pfhip::PostEntry> post;
    std::vector<int> src_list, dft_sim, dft_cell;
    std::vector<pfhip::DftEntry> dft;
    for (int i = 0; i < n; ++i) {
      pf_desc d = descs[i];
      d.flags |= PF_FLAG_GENERIC;
      sims_.emplace_back(new Stepper<real>(d));
      typename Stepper<real>::PostNodes nodes;
      sims_.back()->describe(host[i], pols[i], nodes);
      n_steps_.push_back(host[i].n_steps);
      shapes_.push_back({host[i].nx, host[i].ny});
      pols_.push_back(pols[i]);
      hsrc_.push_back(host[i].src_h.n_groups > 0);
      n_steps_max_ = std::max<int64_t>(n_steps_max_, host[i].n_steps);
      // post nodes of this simulation, one entry per cell in cell order
      std::vector<int> cells;
      for (auto& a : nodes.ade) cells.push_back(a.first);
      for (auto& e : nodes.esrc) cells.push_back(e.first);
      for (auto& e : nodes.dft) cells.push_back(e.cell);
      std::sort(cells.begin(), cells.end());
      cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
      std::sort(nodes.esrc.begin(), nodes.esrc.end());
      std::stable_sort(nodes.dft.begin(), nodes.dft.end(),
                       [](const auto& x, const auto& y) { return x.cell < y.cell; });
      std::sort(nodes.ade.begin(), nodes.ade.end());
      size_t ia = 0, is = 0, id = 0;
      for (int cell : cells) {
        pfhip::PostEntry e{i, cell, -1, 0, 0, 0, 0};
        if (ia < nodes.ade.size() && nodes.ade[ia].first == cell) e.ade_code = nodes.ade[ia++].second;
        e.src0 = (int)src_list.size();
        for (; is < nodes.esrc.size() && nodes.esrc[is].first == cell; ++is)
          src_list.push_back(nodes.esrc[is].second);
        e.src1 = (int)src_list.size();
        e.dft0 = (int)dft.size();
        for (; id < nodes.dft.size() && nodes.dft[id].cell == cell; ++id) {
          const auto& q = nodes.dft[id];
          dft.push_back({q.mon, q.slot, q.comp, q.r});
          dft_sim.push_back(i);
          dft_cell.push_back(cell);
        }
        e.dft1 = (int)dft.size();
        post.push_back(e);
      }
    }
    if (post.size() > (size_t)INT32_MAX / 2 || dft.size() > (size_t)INT32_MAX / 64)
      throw ArgError{"too many post-pass nodes in a batch"};
    sims_dev_ = arena_.upload(host.data(), host.size(), stream_);
    args_.post = post.empty() ? nullptr : arena_.upload(post.data(), post.size(), stream_);
    args_.n_post = (int)post.size();
    args_.src_list = src_list.empty() ? nullptr : arena_.upload(src_list.data(), src_list.size(), stream_);
    args_.dft = dft.empty() ? nullptr : arena_.upload(dft.data(), dft.size(), stream_);
    args_.n_dft = (int)dft.size();
    args_.dft_sim = dft.empty() ? nullptr : arena_.upload(dft_sim.data(), dft_sim.size(), stream_);
    args_.dft_cell = dft.empty() ? nullptr : arena_.upload(dft_cell.data(), dft_cell.size(), stream_);
    args_.ring = dft.empty() ? nullptr
                             : arena_.zeros<real>((size_t)pfhip::BATCH_DFT_RING * dft.size(), stream_);
    for (const auto& h : host)
      for (int m = 0; m < pfhip::BATCH_MAX_DFT; ++m) n_freq_ = std::max(n_freq_, h.dft[m].n_freq);
    step_dev_ = arena_.zeros<int64_t>(1, stream_);
    done_ = arena_.zeros<unsigned>(1, stream_);
    bar_ = arena_.zeros<unsigned>(2, stream_);
    fail_ = arena_.zeros<int>(1, stream_);
    int dev = 0, khz = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    wait_ticks_ = (uint64_t)std::max(khz, 1) * 2000;   // 2 s per barrier wait
    HIP_CHECK(hipEventCreate(&ev_[0]));
    HIP_CHECK(hipEventCreate(&ev_[1]));
    configure(2, 4, 64);
    HIP_CHECK(hipStreamSynchronize(stream_));
  }
  ~Batch() override {
    drop_graph();
    for (hipEvent_t e : ev_)
      if (e) (void)hipEventDestroy(e);
    if (stream_) {
      (void)hipStreamSynchronize(stream_);
      arena_.release();
      (void)hipStreamDestroy(stream_);
    }
  }
  pf_handle* sim(int i) override {
    if (i < 0 || i >= (int)sims_.size()) throw ArgError{"simulation index out of range"};
    return sims_[i].get();
  }
  int size() const override { return (int)sims_.size(); }
  double last_ms() const override { return last_ms_; }
  void write_ce(int i, const void* ce) override {
    if (i < 0 || i >= (int)sims_.size()) throw ArgError{"simulation index out of range"};
    sims_[i]->write_ce(ce, stream_);
    HIP_CHECK(hipStreamSynchronize(stream_));
  }

  // mode 0: one march launch per step; 1: one persistent launch for all
  // steps of an advance; 2: march launches replayed from a captured graph of
  // BATCH_DFT_RING steps and their DFT fold. rows: owned rows per wavefront chunk.
  void configure(int mode, int waves_per_block, int rows) override {
    if (mode < 0 || mode > 2 || (waves_per_block != 1 && waves_per_block != 2 &&
                                 waves_per_block != 4 && waves_per_block != 8) || rows < 1)
      throw ArgError{"bad batch configuration"};
    drop_graph();
    mode_ = mode;
    wpb_ = waves_per_block;
    if (rows != rows_) {
      // chunks grouped by kernel variant (polarisation, H-side sources)
      rows_ = rows;
      std::vector<pfhip::MarchChunk> cs;
      groups_.clear();
      for (int pol = 0; pol < 2; ++pol)
        for (int hs = 0; hs < 2; ++hs) {
          pfhip::ChunkGroup g{(int)cs.size(), 0, pol, hs};
          for (size_t s = 0; s < sims_.size(); ++s) {
            if (pols_[s] != pol || (int)hsrc_[s] != hs) continue;
            const int nx = shapes_[s][0], ny = shapes_[s][1];
            for (int i0 = 0; i0 < nx; i0 += rows)
              for (int j = 0; j < ny; j += pfhip::MARCH_OWNED)
                cs.push_back({(int)s, j - 1, i0, std::min(i0 + rows, nx)});
          }
          g.count = (int)cs.size() - g.first;
          if (g.count > 0) groups_.push_back(g);
        }
      chunks_ = arena_.upload(cs.data(), cs.size(), stream_);
      n_chunks_ = (int)cs.size();
    }
    args_.sims = sims_dev_;
    args_.chunks = chunks_;
    args_.n_chunks = n_chunks_;
    args_.step = step_dev_;
    args_.done = done_;
    args_.bar = bar_;
    args_.fail = fail_;
    args_.wait_ticks = wait_ticks_;
  }

  void reset() override { reset_all(); }

  // Workgroups of the persistent launch: at most half of what the occupancy
  // calculator says fits at once, so all of them are resident together.
  void persistent_grid(int waves_per_block, int* per_cu, int* n_cu, int* grid) override {
    HIP_CHECK(pfhip::batch_persistent_capacity<real>(exact_, waves_per_block, per_cu, n_cu));
    if (*per_cu < 1 || *n_cu < 1) throw HipError{PF_E_STATE, "persistent kernel does not fit a CU"};
    const int need = (n_chunks_ + waves_per_block - 1) / waves_per_block;
    *grid = std::max(1, std::min(std::max(1, *per_cu * *n_cu / 2), need));
  }

  void advance(int64_t n) override {
    if (n < 0 || step_ + n > n_steps_max_) throw ArgError{"step range exceeds the batch"};
    HIP_CHECK(hipMemcpyAsync(step_dev_, &step_, sizeof(int64_t), hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    HIP_CHECK(hipEventRecord(ev_[0], stream_));
    if (mode_ == 1) {
      int per_cu = 0, n_cu = 0, grid = 0;
      persistent_grid(wpb_, &per_cu, &n_cu, &grid);
      pfhip::launch_batch_sync_reset(bar_, fail_, stream_);
      if (n > 0)
        pfhip::launch_batch_persistent<real>(args_, exact_, wpb_, grid, step_, step_ + n, n_freq_,
                                             stream_);
    } else {
      int64_t done = 0;
      if (mode_ == 2) {
        if (!graph_exec_) build_graph();
        for (; done + pfhip::BATCH_DFT_RING <= n; done += pfhip::BATCH_DFT_RING)
          HIP_CHECK(hipGraphLaunch(graph_exec_, stream_));
      }
      // DFT samples are folded at least every BATCH_DFT_RING steps and at the end
      while (done < n) {
        const int span = (int)std::min<int64_t>(pfhip::BATCH_DFT_RING, n - done);
        for (int k = 0; k < span; ++k) launch_step();
        pfhip::launch_batch_dft_fold<real>(args_, span, n_freq_, stream_);
        done += span;
      }
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ev_[1], stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    float t = 0.f;
    HIP_CHECK(hipEventElapsedTime(&t, ev_[0], ev_[1]));
    last_ms_ = t;
    if (mode_ == 1) {
      int failed = 0;
      copy_sync(&failed, fail_, sizeof(int), hipMemcpyDeviceToHost, stream_);
      if (failed)
        throw HipError{PF_E_STATE, "persistent launch: a grid barrier wait expired (workgroups "
                                   "were not all resident); the batch state is undefined, reset it"};
    }
    step_ += n;
    for (size_t i = 0; i < sims_.size(); ++i)
      sims_[i]->set_level((int)(std::min(step_, n_steps_[i]) & 1));
  }

 private:
  void reset_all() {
    for (auto& s : sims_) s->reset(stream_);
    step_ = 0;
    HIP_CHECK(hipMemsetAsync(step_dev_, 0, sizeof(int64_t), stream_));
    HIP_CHECK(hipMemsetAsync(done_, 0, sizeof(unsigned), stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
  }
  // One step: the march variants, then the post pass (which advances the
  // device step counter).
  void launch_step() {
    for (const pfhip::ChunkGroup& g : groups_)
      pfhip::launch_batch_march<real>(args_, g, exact_, wpb_, stream_);
    pfhip::launch_batch_post<real>(args_, stream_);
  }
  // BATCH_DFT_RING steps and the DFT fold of their samples.
  void build_graph() {
    hipGraph_t g;
    HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < pfhip::BATCH_DFT_RING; ++s) launch_step();
    pfhip::launch_batch_dft_fold<real>(args_, pfhip::BATCH_DFT_RING, n_freq_, stream_);
    HIP_CHECK(hipStreamEndCapture(stream_, &g));
    HIP_CHECK(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphDestroy(g));
  }
  void drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
  }

  std::vector<std::unique_ptr<Stepper<real>>> sims_;
  std::vector<int64_t> n_steps_;
  std::vector<std::array<int, 2>> shapes_;
  std::vector<int> pols_;
  std::vector<bool> hsrc_;
  std::vector<pfhip::ChunkGroup> groups_;
  DeviceArena arena_;
  hipStream_t stream_ = nullptr;
  pfhip::SimDev<real>* sims_dev_ = nullptr;
  pfhip::MarchChunk* chunks_ = nullptr;
  int n_chunks_ = 0, rows_ = 0;
  int64_t* step_dev_ = nullptr;
  unsigned* done_ = nullptr;
  int n_freq_ = 0;
  unsigned* bar_ = nullptr;
  int* fail_ = nullptr;
  uint64_t wait_ticks_ = 0;
  int64_t step_ = 0, n_steps_max_ = 0;
  int mode_ = 2, wpb_ = 4;
  bool exact_;
  hipGraphExec_t graph_exec_ = nullptr;
  pfhip::BatchArgs<real> args_{};
  hipEvent_t ev_[2] = {nullptr, nullptr};
  double last_ms_ = 0.0;
};

// ---------------------------------------------------------------------------
// Batch of 3-D simulations: per step, one march launch covers the psi-free
// tiles and one the CPML tiles of every simulation still running (the tile's
// pad0 field names its simulation, whose march arguments come from a device
// table per time level); then one launch each runs the ADE media, the
// E-side sources and the DFT captures of all of them (job tables; the DFT
// jobs carry each ring's fill as the advance starts, and the host keeps
// the rings' bookkeeping and folds). Simulations are
// ordered by decreasing step count so the running ones own a prefix of the
// tile and job lists; each simulation's tiles are in XCD order, padded to a
// multiple of 8.
// ---------------------------------------------------------------------------
template <typename real>
class Batch3 final : public BatchBase {
 public:
  Batch3(const pf_desc* descs, int n) : exact_(descs[0].exact != 0) {
    if (n < 1) throw ArgError{"empty batch"};
    for (int i = 1; i < n; ++i)
      if ((descs[i].exact != 0) != exact_) throw ArgError{"a batch needs one arithmetic mode"};
    HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (int i = 0; i < n; ++i) {
      pf_desc d = descs[i];
      d.flags &= ~PF_FLAG_GENERIC;
      // padded 62-node blocks for every member: one layout for the whole batch,
      // and the batch runs only the psi-free and all-axis CPML tile lists,
      // where that layout's kernel is the faster one
      sims_.emplace_back(new Stepper<real>(d, 62));
      sims_.back()->check_march_batch();
      sims_.back()->adopt_stream(stream_);
      n_steps_.push_back(d.n_steps);
      n_steps_max_ = std::max<int64_t>(n_steps_max_, d.n_steps);
    }
    dist_ = sims_[0]->march_dist();
    zpad_ = sims_[0]->zpad();
    for (auto& s : sims_) {
      if (s->march_dist() != dist_) throw ArgError{"a batch needs one march prefetch distance"};
      if (s->zpad() != zpad_) throw ArgError{"a batch needs one field row layout"};
    }
    for (int p = 0; p < 2; ++p) {
      std::vector<pfhip::MarchArgs<real>> t;
      for (auto& s : sims_) t.push_back(s->march_args(p));
      table_[p] = arena_.upload(t.data(), t.size(), stream_);
    }
    order_.resize(n);
    for (int i = 0; i < n; ++i) order_[i] = i;
    std::stable_sort(order_.begin(), order_.end(),
                     [&](int a, int b) { return n_steps_[a] > n_steps_[b]; });
    // ADE and source jobs per time level, in order_ (prefix counts per
    // number of running simulations)
    for (int p = 0; p < 2; ++p) {
      std::vector<pfhip::AdeJob<real>> ade;
      std::vector<pfhip::SrcJob<real>> src;
      ade_prefix_[p].assign(1, 0);
      src_prefix_[p].assign(1, 0);
      for (int i : order_) {
        sims_[i]->ade_jobs(p, ade);
        pfhip::SrcJob<real> J;
        if (sims_[i]->src_job(p, J)) {
          src.push_back(J);
          max_groups_ = std::max(max_groups_, J.g.n_groups);
        }
        ade_prefix_[p].push_back((int)ade.size());
        src_prefix_[p].push_back((int)src.size());
      }
      for (const auto& J : ade) max_ade_ = std::max(max_ade_, J.n);
      ade_tab_[p] = ade.empty() ? nullptr : arena_.upload(ade.data(), ade.size(), stream_);
      src_tab_[p] = src.empty() ? nullptr : arena_.upload(src.data(), src.size(), stream_);
    }
    // DFT jobs: their ring state is refreshed at every advance
    {
      std::vector<pfhip::DftJob<real>> dft;
      dft_prefix_.assign(1, 0);
      for (int i : order_) {
        sims_[i]->dft_jobs(1, 0, dft);
        dft_prefix_.push_back((int)dft.size());
      }
      for (const auto& J : dft) max_dft_ = std::max<int64_t>(max_dft_, J.nbox * J.n_comp);
      for (int p = 0; p < 2; ++p)
        dft_tab_[p] = dft.empty() ? nullptr
                                  : static_cast<pfhip::DftJob<real>*>(
                                        arena_.alloc(dft.size() * sizeof(pfhip::DftJob<real>)));
    }
    HIP_CHECK(hipEventCreate(&ev_[0]));
    HIP_CHECK(hipEventCreate(&ev_[1]));
    step_dev_ = arena_.zeros<int64_t>(1, stream_);
    configure(0, pfhip::MARCH_NW, pfhip::MARCH_MAX_XCHUNK);
    HIP_CHECK(hipStreamSynchronize(stream_));
  }
  ~Batch3() override {
    drop_graphs();
    for (hipEvent_t e : ev_)
      if (e) (void)hipEventDestroy(e);
    if (stream_) {
      (void)hipStreamSynchronize(stream_);
      sims_.clear();
      arena_.release();
      (void)hipStreamDestroy(stream_);
    }
  }
  pf_handle* sim(int i) override {
    if (i < 0 || i >= (int)sims_.size()) throw ArgError{"simulation index out of range"};
    return sims_[i].get();
  }
  int size() const override { return (int)sims_.size(); }
  double last_ms() const override { return last_ms_; }
  void write_ce(int i, const void* ce) override {
    if (i < 0 || i >= (int)sims_.size()) throw ArgError{"simulation index out of range"};
    sims_[i]->write_ce(ce, stream_);
    HIP_CHECK(hipStreamSynchronize(stream_));
  }

  // mode: 0 launches per step, 2 pairs of steps replayed from a captured
  // hipGraph (no persistent mode); waves_per_block: wavefronts per march
  // workgroup (10, 12 or 16); rows: x planes per psi-free tile
  // (1 .. MARCH_MAX_XCHUNK).
  void configure(int mode, int waves_per_block, int rows) override {
    if (mode != 0 && mode != 2) throw ArgError{"3-D batches run in stream (0) or graph (2) mode"};
    if ((waves_per_block != 10 && waves_per_block != 12 && waves_per_block != 16) || rows < 1 ||
        rows > pfhip::MARCH_MAX_XCHUNK)
      throw ArgError{"bad 3-D batch configuration: waves 10, 12 or 16; rows 1 .. 128"};
    mode_ = mode;
    if (waves_per_block == nw_ && rows == xchunk_) return;
    HIP_CHECK(hipStreamSynchronize(stream_));
    drop_graphs();   // they hold the old tile lists
    for (int q = 0; q < 2; ++q) {   // q = 0: psi-free tiles, 1: CPML tiles
      std::vector<pfhip::MarchTile> all;
      prefix_[q].assign(1, 0);
      for (int i : order_) {
        std::vector<pfhip::MarchTile> t =
            xcd_order(sims_[i]->march_tile_list(waves_per_block, rows, q == 1));
        for (pfhip::MarchTile& x : t) x.pad0 = i;
        all.insert(all.end(), t.begin(), t.end());
        prefix_[q].push_back(all.size());
      }
      if (all.size() > (size_t)INT32_MAX) throw ArgError{"too many march tiles in a batch"};
      if (tiles_[q]) arena_.free(tiles_[q]);
      tiles_[q] = all.empty() ? nullptr : reinterpret_cast<int*>(arena_.upload(all.data(), all.size(), stream_));
    }
    nw_ = waves_per_block;
    xchunk_ = rows;
  }

  void reset() override {
    for (auto& s : sims_) s->reset(stream_);
    step_ = 0;
    HIP_CHECK(hipStreamSynchronize(stream_));
  }

  void persistent_grid(int, int*, int*, int*) override {
    throw ArgError{"3-D batches have no persistent mode"};
  }

  void advance(int64_t n) override {
    if (n < 0 || step_ + n > n_steps_max_) throw ArgError{"step range exceeds the batch"};
    if (dft_tab_[0]) {
      // the rings' fill as this advance starts, for both sampled levels
      HIP_CHECK(hipStreamSynchronize(stream_));
      for (int p = 0; p < 2; ++p) {
        std::vector<pfhip::DftJob<real>> dft;
        for (int i : order_) sims_[i]->dft_jobs(p ^ 1, step_, dft);
        copy_sync(dft_tab_[p], dft.data(), dft.size() * sizeof(pfhip::DftJob<real>),
                  hipMemcpyHostToDevice, stream_);
      }
    }
    HIP_CHECK(hipEventRecord(ev_[0], stream_));
    int64_t counter = -1;   // the device step counter as the queued work leaves it
    for (int64_t s = step_; s < step_ + n;) {
      const size_t running = running_at(s);
      // a pair of steps from the graph: even first step, the same running
      // set for both, and no DFT ring filling at the first (its fold must
      // come before the second step's capture)
      bool pair = mode_ == 2 && (s & 1) == 0 && s + 2 <= step_ + n && running > 0 &&
                  running_at(s + 1) == running;
      for (size_t r = 0; pair && r < running; ++r) pair = !sims_[order_[r]]->dft_ring_fills(s);
      if (pair) {
        if (counter != s) pfhip::launch_step_set(step_dev_, s, stream_);
        HIP_CHECK(hipGraphLaunch(pair_graph(running), stream_));
        counter = s + 2;
        finish_step(s, running);
        finish_step(s + 1, running);
        s += 2;
      } else {
        enqueue_step(s, running, nullptr);
        finish_step(s, running);
        s += 1;
      }
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ev_[1], stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    float t = 0.f;
    HIP_CHECK(hipEventElapsedTime(&t, ev_[0], ev_[1]));
    last_ms_ = t;
    step_ += n;
  }

 private:
  // simulations with steps left at step s: a prefix of order_
  size_t running_at(int64_t s) const {
    size_t running = 0;
    while (running < order_.size() && n_steps_[order_[running]] > s) ++running;
    return running;
  }

  // The launches of one step of the first `running` simulations of order_:
  // the march lists, then the batched ADE, E-source and DFT-capture launches
  // (the single run's order). With step_dev the step comes from the device
  // counter, which the step then advances (graph capture).
  void enqueue_step(int64_t s, size_t running, int64_t* step_dev) {
    const int p = (int)(s & 1);
    for (int q = 0; q < 2; ++q) {   // psi-free tiles, then CPML tiles
      pfhip::MarchArgs<real> L;
      std::memset(&L, 0, sizeof(L));
      L.tiles = tiles_[q];
      L.ntiles = (int)prefix_[q][running];
      L.sims = table_[p];
      if (L.ntiles > 0) {
        if (zpad_ == pfhip::ZPAD_BLOCK) pfhip::launch_march64<real>(L, exact_, q == 1 ? 7 : 0, nw_, stream_);
        else pfhip::launch_march<real>(L, exact_, q == 1, nw_, dist_, stream_);
      }
    }
    pfhip::launch_ade_jobs<real>(ade_tab_[p], ade_prefix_[p][running], max_ade_, stream_);
    pfhip::launch_src_jobs<real>(src_tab_[p], src_prefix_[p][running], max_groups_, s, stream_, step_dev);
    pfhip::launch_dft_jobs<real>(dft_tab_[p], dft_prefix_[running], max_dft_, s, stream_, step_dev);
    if (step_dev) pfhip::launch_step_tick(step_dev, stream_);
  }

  // Host bookkeeping of a step: level flip and DFT ring fill (a full ring
  // is folded on the stream, after the step's capture).
  void finish_step(int64_t s, size_t running) {
    const int p = (int)(s & 1);
    for (size_t r = 0; r < running; ++r) {
      Stepper<real>& sim = *sims_[order_[r]];
      if (sim.level() != p) throw HipError{PF_E_STATE, "batch simulation out of step"};
      sim.finish_step_dft(s);
    }
  }

  // The graph of an even step and the next one for `running` simulations,
  // captured on first use.
  hipGraphExec_t pair_graph(size_t running) {
    for (auto& g : graphs_)
      if (g.first == running) return g.second;
    hipGraph_t graph = nullptr;
    HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    enqueue_step(0, running, step_dev_);
    enqueue_step(1, running, step_dev_);
    HIP_CHECK(hipStreamEndCapture(stream_, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_CHECK(e);
    graphs_.emplace_back(running, exec);
    return exec;
  }

  void drop_graphs() {
    if (!graphs_.empty() && stream_) (void)hipStreamSynchronize(stream_);
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
    graphs_.clear();
  }

  std::vector<std::unique_ptr<Stepper<real>>> sims_;
  std::vector<int64_t> n_steps_;
  std::vector<int> order_;          // simulations by decreasing step count
  std::vector<size_t> prefix_[2];   // tiles of the first r simulations of order_ (per list)
  DeviceArena arena_;
  hipStream_t stream_ = nullptr;
  pfhip::MarchArgs<real>* table_[2] = {nullptr, nullptr};   // per-simulation args reading level p
  pfhip::AdeJob<real>* ade_tab_[2] = {nullptr, nullptr};    // ADE jobs reading level p, in order_
  pfhip::SrcJob<real>* src_tab_[2] = {nullptr, nullptr};    // source jobs writing level p ^ 1
  pfhip::DftJob<real>* dft_tab_[2] = {nullptr, nullptr};    // DFT jobs sampling level p ^ 1
  std::vector<int> dft_prefix_;                             // jobs of the first r simulations
  int64_t max_dft_ = 0;
  std::vector<int> ade_prefix_[2], src_prefix_[2];          // jobs of the first r simulations
  int64_t max_ade_ = 0;
  int max_groups_ = 0;
  int* tiles_[2] = {nullptr, nullptr};   // psi-free tiles, CPML tiles
  int nw_ = 0, xchunk_ = 0, dist_ = 2;
  int zpad_ = 0;    // row layout of the member simulations (ZPAD_BLOCK: the 64-lane march)
  int64_t step_ = 0, n_steps_max_ = 0;
  bool exact_;
  hipEvent_t ev_[2] = {nullptr, nullptr};
  double last_ms_ = 0.0;
  int mode_ = 0;
  int64_t* step_dev_ = nullptr;   // step counter of graph replays
  std::vector<std::pair<size_t, hipGraphExec_t>> graphs_;   // per running count
};

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return PF_OK;
  } catch (const HipError& e) {
    return fail(e.code, e.msg);
  } catch (const ArgError& e) {
    return fail(PF_E_ARG, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(PF_E_LIMIT, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(PF_E_STATE, e.what());
  }
}

}  // namespace

extern "C" {

int pf_create(const pf_desc* desc, pf_handle** out) {
  if (!desc || !out) return fail(PF_E_ARG, "null argument");
  *out = nullptr;
  return guarded([&] {
    if (desc->flags & PF_FLAG_STREAM) *out = pfhost::make_streamer(*desc);
    else if (desc->dtype == PF_F32) *out = new Stepper<float>(*desc);
    else if (desc->dtype == PF_F64) *out = new Stepper<double>(*desc);
    else throw ArgError{"dtype must be PF_F32 or PF_F64"};
  });
}

int pf_destroy(pf_handle* h) {
  return guarded([&] { delete h; });
}

int pf_advance(pf_handle* h, int64_t step0, int64_t n) {
  if (!h) return fail(PF_E_ARG, "null handle");
  return guarded([&] { h->advance(step0, n); });
}

int pf_read_field(pf_handle* h, int32_t comp, void* dst) {
  if (!h) return fail(PF_E_ARG, "null handle");
  pf_box b;
  for (int a = 0; a < 3; ++a) {
    b.start[a] = 0;
    b.count[a] = h->shape[a];
    b.stride[a] = 1;
  }
  return guarded([&] { h->read_box(comp, b, dst); });
}

int pf_read_box(pf_handle* h, int32_t comp, const pf_box* box, void* dst) {
  if (!h || !box) return fail(PF_E_ARG, "null argument");
  return guarded([&] { h->read_box(comp, *box, dst); });
}

int pf_read_box_q(pf_handle* h, int32_t comp, const pf_box* box, int32_t bits, void* dst,
                  double* scale) {
  if (!h || !box) return fail(PF_E_ARG, "null argument");
  return guarded([&] { h->read_box_q(comp, *box, bits, dst, scale); });
}

int pf_write_field(pf_handle* h, int32_t comp, const void* src) {
  if (!h) return fail(PF_E_ARG, "null handle");
  return guarded([&] { h->write_field(comp, src); });
}

int pf_dft_read(pf_handle* h, int32_t mon, int32_t slot, double* dst) {
  if (!h) return fail(PF_E_ARG, "null handle");
  return guarded([&] { h->dft_read(mon, slot, dst); });
}

int pf_stepper_info(pf_handle* h, pf_info* out) {
  if (!h || !out) return fail(PF_E_ARG, "null argument");
  return guarded([&] { h->info(out); });
}

int pf_field_ptr(pf_handle* h, int32_t comp, pf_view* out) {
  if (!h || !out) return fail(PF_E_ARG, "null argument");
  return guarded([&] { h->field_view(comp, out); });
}

int pf_dft_ptr(pf_handle* h, int32_t mon, pf_view* out) {
  if (!h || !out) return fail(PF_E_ARG, "null argument");
  return guarded([&] { h->dft_view(mon, out); });
}

int pf_dft_flush(pf_handle* h) {
  if (!h) return fail(PF_E_ARG, "null handle");
  return guarded([&] { h->dft_flush(); });
}

int pf_flux_read(pf_handle* h, double* dst) {
  if (!h) return fail(PF_E_ARG, "null handle");
  return guarded([&] { h->flux_read(dst); });
}

int pf_timing_enable(pf_handle* h, int32_t on) {
  if (!h) return fail(PF_E_ARG, "null handle");
  if (on < 0) return fail(PF_E_ARG, "negative timing interval");
  return guarded([&] { h->timing_enable(on); });
}

int pf_timing_read(pf_handle* h, double* ms, int64_t* launches) {
  if (!h) return fail(PF_E_ARG, "null handle");
  return guarded([&] { h->timing_read(ms, launches); });
}

int pf_checkpoint_enable(pf_handle* h, int32_t interval) {
  if (!h) return fail(PF_E_ARG, "null handle");
  return guarded([&] { h->checkpoint_enable(interval); });
}

int pf_adjoint(pf_handle* h, int32_t n_cot, const pf_cotangent* cot, double* grad_ce) {
  if (!h) return fail(PF_E_ARG, "null handle");
  if (n_cot < 0 || (n_cot > 0 && !cot)) return fail(PF_E_ARG, "bad cotangent list");
  return guarded([&] { h->adjoint(n_cot, cot, grad_ce); });
}

struct pf_batch {
  std::unique_ptr<BatchBase> impl;
};

int pf_batch_create(const pf_desc* descs, const int32_t* pols, int32_t n, pf_batch** out) {
  if (!descs || !pols || !out || n < 1) return fail(PF_E_ARG, "bad batch arguments");
  *out = nullptr;
  return guarded([&] {
    for (int i = 1; i < n; ++i)
      if (descs[i].dtype != descs[0].dtype) throw ArgError{"a batch needs one dtype"};
    for (int i = 0; i < n; ++i)
      if (descs[i].ce_y || descs[i].ce_z) throw ArgError{"subpixel smoothing is not supported in a batch"};
    int n3 = 0;   // 3-D grids (nz >= 2)
    for (int i = 0; i < n; ++i) n3 += descs[i].shape[2] > 1;
    if (n3 != 0 && n3 != n) throw ArgError{"a batch holds 2-D grids only or 3-D grids only"};
    std::unique_ptr<pf_batch> b(new pf_batch);
    if (descs[0].dtype == PF_F32) {
      if (n3) b->impl.reset(new Batch3<float>(descs, n));
      else b->impl.reset(new Batch<float>(descs, n, pols));
    } else if (descs[0].dtype == PF_F64) {
      if (n3) b->impl.reset(new Batch3<double>(descs, n));
      else b->impl.reset(new Batch<double>(descs, n, pols));
    } else {
      throw ArgError{"dtype must be PF_F32 or PF_F64"};
    }
    *out = b.release();
  });
}

int pf_batch_destroy(pf_batch* b) {
  return guarded([&] { delete b; });
}

int pf_batch_configure(pf_batch* b, int32_t mode, int32_t waves_per_block, int32_t rows) {
  if (!b) return fail(PF_E_ARG, "null batch");
  return guarded([&] { b->impl->configure(mode, waves_per_block, rows); });
}

int pf_batch_reset(pf_batch* b) {
  if (!b) return fail(PF_E_ARG, "null batch");
  return guarded([&] { b->impl->reset(); });
}

int pf_batch_persistent_grid(pf_batch* b, int32_t waves_per_block, int32_t* per_cu, int32_t* n_cu,
                             int32_t* grid) {
  if (!b || !per_cu || !n_cu || !grid) return fail(PF_E_ARG, "null argument");
  return guarded([&] { b->impl->persistent_grid(waves_per_block, per_cu, n_cu, grid); });
}

int pf_batch_advance(pf_batch* b, int64_t n) {
  if (!b) return fail(PF_E_ARG, "null batch");
  return guarded([&] { b->impl->advance(n); });
}

int pf_batch_last_ms(pf_batch* b, double* ms) {
  if (!b || !ms) return fail(PF_E_ARG, "null argument");
  return guarded([&] { *ms = b->impl->last_ms(); });
}

int pf_batch_write_ce(pf_batch* b, int32_t sim, const void* ce) {
  if (!b || !ce) return fail(PF_E_ARG, "null argument");
  return guarded([&] { b->impl->write_ce(sim, ce); });
}

int pf_batch_dft_read(pf_batch* b, int32_t sim, int32_t mon, int32_t slot, double* dst) {
  if (!b) return fail(PF_E_ARG, "null batch");
  return guarded([&] { b->impl->sim(sim)->dft_read(mon, slot, dst); });
}

int pf_batch_read_field(pf_batch* b, int32_t sim, int32_t comp, void* dst) {
  if (!b) return fail(PF_E_ARG, "null batch");
  return guarded([&] {
    pf_handle* h = b->impl->sim(sim);
    pf_box box;
    for (int a = 0; a < 3; ++a) {
      box.start[a] = 0;
      box.count[a] = h->shape[a];
      box.stride[a] = 1;
    }
    h->read_box(comp, box, dst);
  });
}

int pf_device_info(char* name, int32_t name_len, int64_t* free_mb, int64_t* total_mb) {
  return guarded([&] {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    size_t fr = 0, tot = 0;
    HIP_CHECK(hipMemGetInfo(&fr, &tot));
    if (name && name_len > 0) {
      std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (free_mb) *free_mb = (int64_t)(fr >> 20);
    if (total_mb) *total_mb = (int64_t)(tot >> 20);
  });
}

const char* pf_last_error(void) { return g_last_error.c_str(); }

int32_t pf_abi_version(void) { return PFHIP_ABI_VERSION; }

}  // extern "C"
