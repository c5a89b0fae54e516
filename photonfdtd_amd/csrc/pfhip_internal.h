// Internal declarations shared by the libpfhip translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfhip {

// psi slots: E terms then H terms, in (component, +term, -term) order.
enum PsiSlot {
  PSI_EXY = 0, PSI_EXZ, PSI_EYZ, PSI_EYX, PSI_EZX, PSI_EZY,
  PSI_HXY, PSI_HXZ, PSI_HYZ, PSI_HYX, PSI_HZX, PSI_HZY,
  PSI_COUNT
};

// Kernel arguments of one Yee pass, passed by value.
template <typename real>
struct StepArgs {
  real* f[6];            // Ex, Ey, Ez, Hx, Hy, Hz
  const real* ce;        // per-node E coefficient (of Ex when ce_y / ce_z are set)
  const real* ce_y;      // subpixel smoothing: Ey / Ez coefficients, or null
  const real* ce_z;
  real* psi[PSI_COUNT];  // compact CPML state
  // Out-of-place update (null: in place). The H pass reads E and H from f
  // and writes H to fo[3..5]; the E pass reads the new H from fo[3..5] and E
  // from f and writes E to fo[0..2]; psi is read from psi, written to psio.
  real* fo[6];
  real* psio[PSI_COUNT];
  const real* b_e[3];
  const real* c_e[3];
  const real* b_h[3];
  const real* c_h[3];
  const int* map_e[3];   // compact index of each node along the axis, -1 outside PML
  const int* map_h[3];
  int pe[3], ph[3];      // compact lengths
  int nx, ny, nz;
  int64_t nyz;
  real d[3];
  real inv_d[3];
  real ch;
};

// Kernel arguments of the fused ping-pong step (3-D grids).
template <typename real>
struct FusedArgs {
  const real* Ein[3];
  real* Eout[3];
  const real* Hin[3];
  real* Hout[3];
  const real* ce;
  real* psiE[6];          // E-side psi, updated in place
  const real* psiHin[6];  // H-side psi, ping-ponged
  real* psiHout[6];
  const real* b_e[3];
  const real* c_e[3];
  const real* b_h[3];
  const real* c_h[3];
  const int* map_e[3];
  const int* map_h[3];
  int pe[3], ph[3];
  int nx, ny, nz;
  int64_t nyz;
  real d[3];
  real inv_d[3];
  real ch;
  int xchunk;             // x planes per workgroup
  int variant;            // 0: register march, 1: LDS-staged march, 2: march2
  int slab_e[3][2];       // per axis: E-side CPML nodes are n < lo or n >= hi
  int slab_h[3][2];
  int tiles_z, tiles_y, chunks_x, grid_blocks;   // filled by the launcher
};

// Padded z rows of the march layout: with zpad = P nodes per block (62 for
// the single-step march, 60 when the two-step march shares the levels),
// node k of a row sits in slot k + (64 - P) * (k / P), so each P-node z
// tile of the marches owns an aligned 64-slot (256-byte in float32) block
// and its row stores are whole cache lines; the pad slots of a block hold
// no node. zpad = 0: dense rows.
constexpr int ZPAD_BLOCK = 64;
__host__ __device__ inline int64_t zslot(int64_t k, int zpad) {
  return zpad ? k + (ZPAD_BLOCK - zpad) * (k / zpad) : k;
}
// Row length (slots) of the padded layout for nz nodes, zpad nodes per block.
inline int zpad_row(int nz, int zpad) { return ZPAD_BLOCK * ((nz + zpad - 1) / zpad); }

// (j, k) of the flat in-plane index q of a plane of nyz nodes with rows of nz
// (the per-node kernels' thread mapping); 32-bit division when the plane
// allows it.
__device__ __forceinline__ void plane_jk(int64_t q, int64_t nyz, int nz, int& j, int& k) {
  if (nyz <= 0x7FFFFFFF) {
    const uint32_t u = (uint32_t)q, z = (uint32_t)nz;
    const uint32_t jj = u / z;
    j = (int)jj;
    k = (int)(u - jj * z);
  } else {
    j = (int)(q / nz);
    k = (int)(q - (int64_t)j * nz);
  }
}

// Arguments of the march kernel: a few base pointers, everything else derived.
constexpr int MARCH_MAX_XCHUNK = 128;
template <typename real>
struct MarchArgs {
  const real* Lin;        // level read: Ex Ey Ez Hx Hy Hz, each nx*ny*nz
  real* Lout;             // level written
  const real* ce;
  real* psiE;             // E-side psi slots ExY ExZ EyZ EyX EzX EzY, back to back
  const real* psiHin;     // H-side psi slots HxY HxZ HyZ HyX HzX HzY (read)
  real* psiHout;          // H-side psi (written)
  const real* coef;       // [axis][b_e, c_e, b_h, c_h][node]
  int nx, ny, nz;
  int pe[3], ph[3];
  int slab_e[3][2], slab_h[3][2];
  real d[3], inv_d[3], ch;
  int xchunk;
  int zrow, zpad;         // field/ce row length in slots; nodes per padded block (0: dense)
  const int* tiles;       // [ntiles][8]: i0, i1, j0, jn, k0, kn, sim, - (host-built)
  int ntiles;
  // batched launch: per-simulation arguments indexed by the tile's sim field
  // (nullptr for a single simulation)
  const MarchArgs* sims;
};

// Tile of the march kernel: planes [i0, i1), owned rows j0..j0+jn-1 and owned
// z nodes k0..k0+kn-1 (jn <= NW-2, kn <= 62). i1 <= i0 marks padding. pad0
// is the simulation index in a batched launch.
struct MarchTile {
  int i0, i1, j0, jn, k0, kn, pad0, pad1;
};

// ---- two-step march (temporal blocking) ----------------------------------
// Workgroups of NW waves; tile as MarchTile with jn <= NW-4 owned rows,
// kn <= 60 owned z nodes and at most MARCH_MAX_XCHUNK planes.
constexpr int MARCH2_OWNED_Z = 60;
constexpr int MARCH2_MAX_ADE = 4;
constexpr int MARCH2_MAX_DFT = 4;

// A dispersive medium whose cells fill the box [lo, hi) in C order, so the
// cell index t follows from (i, j, k). P arrays are [pole][comp][t].
template <typename real>
struct March2Ade {
  int lo[3], hi[3];
  int n_poles;
  int64_t n;
  const real* abc;
  real inv_eps;
  const real* Pcur_in;    // P^n
  const real* Pprev_in;   // P^{n-1}
  real* Pcur_out;         // P^{n+2}
  real* Pprev_out;        // P^{n+1}
};

// Capture of the first of the two steps into a DFT ring slot: slot s is
// ring + s * slot_elems, laid out [c][box index] for component comps[c].
template <typename real>
struct March2Dft {
  int n_comp;
  int comps[6];
  int start[3], count[3], stride[3];
  real* ring;
  int64_t slot_elems;
};

template <typename real>
struct MarchArgs2 {
  const real* Lin;        // level read (n): Ex Ey Ez Hx Hy Hz
  real* Lout;             // level written (n + 2)
  const real* ce;
  const real* psiEin;     // E-side psi, level read
  real* psiEout;          // E-side psi, level written
  const real* psiHin;
  real* psiHout;
  const real* coef;       // [axis][b_e, c_e, b_h, c_h][node]
  int nx, ny, nz;
  int pe[3], ph[3];
  int slab_e[3][2], slab_h[3][2];
  real d[3], inv_d[3], ch;
  int zrow, zpad;         // field/ce row length in slots; nodes per padded block (0: dense)
  const int* tiles;
  int ntiles;
  int64_t step;           // n: the pass advances n -> n + 2
  int64_t n_steps;        // waveform table length
  // E-side soft sources: groups sorted by (component, cell)
  int src_beg[4];         // groups of component c: [src_beg[c], src_beg[c + 1])
  int src_i_lo, src_i_hi; // x range of the source cells
  const int64_t* src_cell;
  const int32_t* src_begin;
  const double* src_amp;
  const int32_t* src_row;
  const double* src_wave;
  // media and monitors live in device memory and are read only where a
  // node needs them, so they take no registers in the plane loop
  int n_ade;
  const March2Ade<real>* ade;   // [n_ade], P levels of this pass
  int n_dft;
  const March2Dft<real>* dft;   // [n_dft]
  uint32_t dft_slots;           // ring slot of monitor m in bits 8m..8m+7; 0xFF: no sample
};

// pml: the tiles touch CPML slabs; nw: 10 or 12 waves per workgroup.
template <typename real>
void launch_march2(const MarchArgs2<real>& a, bool exact, bool pml, int nw, hipStream_t s);

// Arguments of the adjoint kernels (generic in-place layout).
template <typename real>
struct AdjArgs {
  const real* f[6];          // forward fields: E^n and H^{n+1/2}
  const real* ce;
  const real* psi_fwd[12];   // forward psi before the respective update
  real* lam_f[6];            // adjoint fields (updated in place)
  real* lam_psi[12];         // adjoint psi (updated in place)
  real* lam_d[6];            // per-term scratch of the E update, one nx*ny*nz array each
  real* lam_dh[6];           // per-term scratch of the H update
  double* grad;              // dJ/dce accumulator
  const real* b_e[3];
  const real* c_e[3];
  const real* b_h[3];
  const real* c_h[3];
  const int* map_e[3];
  const int* map_h[3];
  int pe[3], ph[3];
  int nx, ny, nz;
  int64_t nyz;
  real inv_d[3];
  real ch;
};

struct BoxDev {
  int64_t start[3];
  int64_t count[3];
  int64_t stride[3];
};

struct SourceGroupsDev {
  int32_t n_groups = 0;
  int32_t* comp = nullptr;       // [n_groups]
  int64_t* cell = nullptr;       // [n_groups]
  int32_t* begin = nullptr;      // [n_groups + 1] into amp / wave_row
  double* amp = nullptr;         // [n_src]
  int32_t* wave_row = nullptr;   // [n_src]
  double* wave = nullptr;        // [n_wave][n_steps]
};

// Soft sources of group g: summed in source order, rounded to the field
// dtype after every add (the per-dipole `F += amp * w(t)` of the reference
// loop, simulation.py:1366-1390).
template <typename real>
__device__ inline void inject_group(real* const* fields, const SourceGroupsDev& G, int64_t n_steps,
                                    int64_t step, int g) {
  real* F = fields[G.comp[g]];
  const int64_t cell = G.cell[g];
  real v = F[cell];
  for (int s = G.begin[g]; s < G.begin[g + 1]; ++s) {
    const double add = G.amp[s] * G.wave[(int64_t)G.wave_row[s] * n_steps + step];
    v = (real)((double)v + add);
  }
  F[cell] = v;
}

// ADE correction of node t of a dispersive medium for the ping-pong layout:
// E^n from Ein, the curl-updated E^{n+1} corrected in place in Eout; P is
// [pole][comp][t], updated in place (simulation.py:638-669).
template <typename real>
__device__ inline void ade_pp_node(const real* const* Ein, real* const* Eout, const int64_t* cells,
                                   int64_t n, real* Pcur, real* Pprev, const real* abc, int n_poles,
                                   real inv_eps, int64_t t) {
  const int64_t cell = cells[t];
  for (int c = 0; c < 3; ++c) {
    const real e = Ein[c][cell];
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
    Eout[c][cell] = Eout[c][cell] - inv_eps * delta;
  }
}

// Batched 3-D stepping: the ADE media and E-side sources of all simulations
// of a batch in one launch each. One job per (simulation, medium) and per
// simulation, for one time level; a launch takes the jobs of the running
// simulations (a prefix of the table).
template <typename real>
struct AdeJob {
  const real* Ein[3];
  real* Eout[3];
  const int64_t* cells;
  int64_t n;
  real* Pcur;
  real* Pprev;
  const real* abc;
  int n_poles;
  real inv_eps;
};
template <typename real>
struct SrcJob {
  real* const* fields;    // device table of the output level's six components
  SourceGroupsDev g;
  int64_t n_steps;
};
// One DFT monitor of one simulation: the step's sample goes to ring slot
// (count0 + q - q0) mod K for sample q = step / interval (steps off the
// interval take none); q0 / count0: the first sample of the advance and the
// ring's fill when it started.
template <typename real>
struct DftJob {
  real* const* fields;    // device table of the sampled level's six components
  const int32_t* comps;
  int n_comp;
  BoxDev box;
  int64_t plane;
  int row, zpad;
  real* ring;
  int64_t nbox;
  int K, interval;
  int64_t q0;
  int count0;
};
template <typename real>
void launch_ade_jobs(const AdeJob<real>* jobs, int n_jobs, int64_t max_n, hipStream_t s);
// One step's ADE media, E-side sources and DFT samples in one launch
// (k_post_step), for a single simulation whose parts touch disjoint E nodes.
constexpr int POST_MAX_ADE = 8;
constexpr int POST_MAX_DFT = 8;
template <typename real>
struct PostGather {
  const int32_t* comps;   // device array of n_comp component codes
  BoxDev box;
  int64_t nbox;
  int blocks_c;           // blocks per component: ceil(nbox / 256)
  real* out;              // ring slot [n_comp][nbox]
};
template <typename real>
struct PostArgs {
  const AdeJob<real>* ade;   // device table of n_ade jobs
  int n_ade;
  int ade_blk[POST_MAX_ADE + 1];
  real* const* src_fields;   // device table of the output level
  SourceGroupsDev src;
  int64_t n_steps, step;
  int src_blk;
  real* const* dft_fields;
  int64_t plane;
  int row, zpad;
  int n_dft;
  int dft_blk[POST_MAX_DFT + 1];
  PostGather<real> dft[POST_MAX_DFT];
};
template <typename real>
void launch_post_step(const PostArgs<real>& P, hipStream_t s);

// The step is `step`, or *step_dev when step_dev is set (graph replays).
template <typename real>
void launch_dft_jobs(const DftJob<real>* jobs, int n_jobs, int64_t max_elems, int64_t step,
                     hipStream_t s, const int64_t* step_dev = nullptr);
template <typename real>
void launch_src_jobs(const SrcJob<real>* jobs, int n_jobs, int max_groups, int64_t step, hipStream_t s,
                     const int64_t* step_dev = nullptr);
// Device step counter of graph replays: += 1, = v.
void launch_step_tick(int64_t* step, hipStream_t s);
void launch_step_set(int64_t* step, int64_t v, hipStream_t s);

// ---- batched 2-D stepping -------------------------------------------------
struct DftDev {
  int n_comp;
  int comps[6];
  int64_t nbox;
  int n_freq, interval;
  const double* phase_e;
  const double* phase_h;
  double dt;
  double* acc;            // [slot][freq][box][2]
};

template <typename real>
struct AdeDev {
  int64_t n;              // cells of this medium
  int n_poles;
  const real* abc;
  real inv_eps;
  real* Pcur;
  real* Pprev;
};

constexpr int BATCH_MAX_DFT = 8;
constexpr int BATCH_MAX_ADE = 8;

// One simulation of a batch: an xy-plane grid (nz = 1) with two time levels
// of fields and of the H-side CPML psi (the march reads level p and writes
// level p ^ 1, p = step & 1); E-side psi and ADE state are updated in place.
template <typename real>
struct SimDev {
  real* f[2][6];
  real* psiH[2][6];       // H-side psi slots (PSI_HXY - 6 ...) per level
  real* psiE[6];          // E-side psi slots (PSI_EXY ...)
  const real* ce;
  const real* b_e[2];
  const real* c_e[2];
  const real* b_h[2];
  const real* c_h[2];
  int slab_e[2][2], slab_h[2][2];   // CPML nodes of x / y: [0, lo) U [hi, n)
  int pe[2], ph[2];                 // compact CPML extent per axis
  int nx, ny;
  int nyr;                // field / ce row length in slots (padded rows, zslot(j, MARCH_OWNED))
  real d[2], inv_d[2], ch;
  int pol;                // 0: TE (Ex, Ey, Hz), 1: TM (Ez, Hx, Hy)
  int64_t n_steps;
  SourceGroupsDev src_e, src_h;
  int64_t src_lo, src_hi; // range of H-side source cells
  AdeDev<real> ade[BATCH_MAX_ADE];
  DftDev dft[BATCH_MAX_DFT];
};

// Work of one wavefront: lane l covers column j0 + l (lanes 1..62 are owned,
// lanes 0 and 63 only supply neighbours), rows [i0, i1) are owned and row
// i0 - 1 is recomputed for its H.
struct MarchChunk {
  int sim, j0, i0, i1;
};
constexpr int MARCH_OWNED = 62;

// A node that needs work after the curl update, in the reference's order:
// ADE correction (code >= 0: (medium << 26) | index), E-side soft sources
// (source groups [src0, src1) of src_list), DFT samples (dft entries
// [dft0, dft1)).
struct PostEntry {
  int sim, cell, ade_code;
  int src0, src1, dft0, dft1;
};

// One DFT sample stream: component `comp` of monitor `mon` at box index r.
struct DftEntry {
  int mon, slot, comp;
  int64_t r;
};

constexpr int BATCH_DFT_RING = 32;   // steps of DFT samples buffered before a fold

template <typename real>
struct BatchArgs {
  const SimDev<real>* sims;
  const MarchChunk* chunks;
  int n_chunks;
  const PostEntry* post;
  int n_post;
  const int* src_list;
  const DftEntry* dft;
  int n_dft;
  const int* dft_sim;      // simulation of each DFT entry
  const int* dft_cell;     // cell of each DFT entry
  real* ring;              // [BATCH_DFT_RING][n_dft]
  int64_t* step;           // device step counter (per-step launches)
  unsigned* done;          // workgroups of the post pass that finished this step
  unsigned* bar;           // persistent mode: [arrived, generation]
  int* fail;               // persistent mode: set when a barrier wait expired
  uint64_t wait_ticks;     // persistent mode: bound of one barrier wait (wall clock)
};

// Chunks [first, first + count) of one kernel variant: polarisation and
// whether the simulation has H-side sources.
struct ChunkGroup {
  int first, count, pol, hsrc;
};

template <typename real>
void launch_batch_march(const BatchArgs<real>& a, const ChunkGroup& g, bool exact,
                        int waves_per_block, hipStream_t s);
template <typename real>
void launch_batch_post(const BatchArgs<real>& a, hipStream_t s);
template <typename real>
void launch_batch_dft_fold(const BatchArgs<real>& a, int span, int n_freq, hipStream_t s);
void launch_batch_sync_reset(unsigned* bar, int* fail, hipStream_t s);
template <typename real>
void launch_batch_persistent(const BatchArgs<real>& a, bool exact, int waves_per_block, int blocks,
                             int64_t step0, int64_t step1, int n_freq, hipStream_t s);
template <typename real>
hipError_t batch_persistent_capacity(bool exact, int waves_per_block, int* per_cu, int* n_cu);

// Snapshot quantisation for compressed FieldMonitor storage: scale = max |x|
// (1 when 0 or not finite) into *scale, x / scale * 127 rounded half to even
// into int8 (bits 8) or x / scale rounded to float16 bits (bits 16). part
// holds max_part doubles of partial maxima.
template <typename real>
void launch_quantize(const real* x, int64_t n, int bits, double* part, int max_part, void* out,
                     double* scale, hipStream_t s);
template <typename real> void launch_generic_h(const StepArgs<real>& a, bool exact, hipStream_t s);
template <typename real> void launch_generic_e(const StepArgs<real>& a, bool exact, hipStream_t s);
template <typename real>
void launch_inject(real* const* fields_dev, const SourceGroupsDev& g, int64_t n_steps, int64_t step,
                   hipStream_t s);
template <typename real>
void launch_ade_capture(real* const* fields_dev, const int64_t* cells, int64_t n, real* En,
                        hipStream_t s);
template <typename real>
void launch_ade_apply(real* const* fields_dev, const int64_t* cells, int64_t n, const real* En,
                      real* Pcur, real* Pprev, const real* abc, int n_poles, real inv_eps,
                      hipStream_t s);
// plane / row: field strides in elements (ny * row, row); zpad: padded rows.
template <typename real>
void launch_gather_box(real* const* fields_dev, const int32_t* comps, int n_comp, const BoxDev& box,
                       int64_t plane, int row, int zpad, real* out, hipStream_t s,
                       int64_t cstride = 0);
template <typename real>
void launch_dft_fold(const real* ring, int64_t nbox, int n_comp, const int32_t* is_h, int count,
                     int64_t sample0, const double* phase_e, const double* phase_h, int n_freq,
                     double dt, double* acc, hipStream_t s);
template <typename real>
void launch_flux(real* const* fields_host, int axis, int64_t index, int nx, int ny, int nz,
                 int zrow, int zpad, double dA, double dt, double* partial, int n_part,
                 double* total, hipStream_t s);

constexpr int MARCH_NW = 10;   // wavefronts per march workgroup (rows j0-1 .. j0+NW-2)
template <typename real>
void launch_march(MarchArgs<real>& a, bool exact, bool pml, int nw, int dist, hipStream_t s);
// The march on dense rows (zpad == ZPAD_BLOCK: node k in slot k) with tiles of
// up to 64 z nodes starting on 64-node boundaries (k_yee_m64). mask: the CPML
// axes of the tiles (bit 0 x, 1 y, 2 z): 0 and 7 for any launch, 1, 2 and 4
// for single-simulation launches.
template <typename real>
void launch_march64(MarchArgs<real>& a, bool exact, int mask, int nw, hipStream_t s);
// Tiles of the psi-free and z-only CPML sets in one launch, with the y-only
// (sets bit 1) and x-only (sets bit 0; with y only) sets, the set of each
// tile in its pad1 field (0, 4, 2, 1).
template <typename real>
void launch_march64_mix(MarchArgs<real>& a, bool exact, int nw, int sets, hipStream_t s);
template <typename real> bool fused_supported(const StepArgs<real>& a);
template <typename real> void launch_fused(FusedArgs<real>& a, bool exact, hipStream_t s);
template <typename real>
void launch_ade_pp(real* const* Ein, real* const* Eout, const int64_t* cells, int64_t n, real* Pcur,
                   real* Pprev, const real* abc, int n_poles, real inv_eps, hipStream_t s);

// phase 0: transpose of the E update, after (gather) adding the previous
// reverse step's H-update terms to lam_E; phase 1: the E-update terms added
// to lam_H, then the transpose of the H update.
template <typename real>
void launch_adjoint_step(const AdjArgs<real>& a, hipStream_t s, int phase, bool gather);
template <typename real>
void launch_adj_dft_source(real* const* lam_dev, const int32_t* comps, const int32_t* is_h,
                           int n_comp, const BoxDev& box, int64_t nyz, int nz, const double* cot,
                           int n_freq, const double* phase_e, const double* phase_h,
                           int64_t sample, double dt, hipStream_t s);

}  // namespace pfhip
