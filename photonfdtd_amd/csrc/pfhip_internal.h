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
  const real* ce;        // per-node E coefficient
  real* psi[PSI_COUNT];  // compact CPML state
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
  const int* tiles;       // [ntiles][8]: i0, i1, j0, jn, k0, kn, -, - (host-built)
  int ntiles;
};

// Tile of the march kernel: planes [i0, i1), owned rows j0..j0+jn-1 and owned
// z nodes k0..k0+kn-1 (jn <= NW-2, kn <= 62). i1 <= i0 marks padding.
struct MarchTile {
  int i0, i1, j0, jn, k0, kn, pad0, pad1;
};

// Arguments of the adjoint kernels (generic in-place layout).
template <typename real>
struct AdjArgs {
  const real* f[6];          // forward fields: E^n and H^{n+1/2}
  const real* ce;
  const real* psi_fwd[12];   // forward psi before the respective update
  real* lam_f[6];            // adjoint fields (updated in place)
  real* lam_psi[12];         // adjoint psi (updated in place)
  real* lam_d[6];            // per-term scratch, one nx*ny*nz array each
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

// ---- batched 2-D stepping -------------------------------------------------
struct DftDev {
  int n_comp;
  int comps[6];
  BoxDev box;
  int lo[2], hi[2];       // first and last (i, j) node of the box
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

// One simulation of a batch: an xy-plane grid (nz = 1), in-place fields.
template <typename real>
struct SimDev {
  real* f[6];
  const real* ce;
  real* psi[PSI_COUNT];
  const real* b_e[3];
  const real* c_e[3];
  const real* b_h[3];
  const real* c_h[3];
  const int* map_e[3];
  const int* map_h[3];
  int pe[3], ph[3];
  int nx, ny;
  real d[3], inv_d[3], ch;
  int pol;                // 0: TE (Ex, Ey, Hz), 1: TM (Ez, Hx, Hy)
  int64_t n_steps;
  SourceGroupsDev src_e, src_h;
  int64_t src_lo[2], src_hi[2];   // range of source cells per side (0: E, 1: H)
  const int* ade_code;    // per node: (medium << 26) | index, or -1
  int n_ade;
  AdeDev<real> ade[BATCH_MAX_ADE];
  int n_dft;
  DftDev dft[BATCH_MAX_DFT];
};

template <typename real>
struct BatchArgs {
  const SimDev<real>* sims;
  const int* chunk_sim;
  const int* chunk_first;
  int n_chunks;
  int chunk_cells;
  const int64_t* step;    // device step counter (per-phase kernels)
};

template <typename real>
void launch_batch_phase(const BatchArgs<real>& a, int phase, bool exact, int grid, int block,
                        hipStream_t s);
void launch_batch_tick(int64_t* step, hipStream_t s);
template <typename real>
hipError_t launch_batch_persistent(BatchArgs<real>& a, bool exact, int block, int64_t step0,
                                   int64_t step1, hipStream_t s);
template <typename real>
hipError_t batch_persistent_grid(bool exact, int block, int n_chunks, int* per_cu, int* n_cu,
                                 int* grid);

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
template <typename real>
void launch_gather_box(real* const* fields_dev, const int32_t* comps, int n_comp, const BoxDev& box,
                       int64_t nyz, int nz, real* out, hipStream_t s);
template <typename real>
void launch_dft_fold(const real* ring, int64_t nbox, int n_comp, const int32_t* is_h, int count,
                     int64_t sample0, const double* phase_e, const double* phase_h, int n_freq,
                     double dt, double* acc, hipStream_t s);
template <typename real>
void launch_flux(real* const* fields_host, int axis, int64_t index, int nx, int ny, int nz,
                 double dA, double dt, double* partial, int n_part, double* total, hipStream_t s);

constexpr int MARCH_NW = 10;   // wavefronts per march workgroup (rows j0-1 .. j0+NW-2)
template <typename real>
void launch_march(MarchArgs<real>& a, bool exact, bool pml, int nw, hipStream_t s);
template <typename real> bool fused_supported(const StepArgs<real>& a);
template <typename real> void launch_fused(FusedArgs<real>& a, bool exact, hipStream_t s);
template <typename real>
void launch_ade_pp(real* const* Ein, real* const* Eout, const int64_t* cells, int64_t n, real* Pcur,
                   real* Pprev, const real* abc, int n_poles, real inv_eps, hipStream_t s);

template <typename real> void launch_adjoint_step(const AdjArgs<real>& a, hipStream_t s, int phase);
template <typename real>
void launch_adj_dft_source(real* const* lam_dev, const int32_t* comps, const int32_t* is_h,
                           int n_comp, const BoxDev& box, int64_t nyz, int nz, const double* cot,
                           int n_freq, const double* phase_e, const double* phase_h,
                           int64_t sample, double dt, hipStream_t s);

}  // namespace pfhip
