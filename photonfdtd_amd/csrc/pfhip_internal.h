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

}  // namespace pfhip
