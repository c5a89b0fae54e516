/*
 * pfhip.h -- C ABI of libpfhip.so, the gfx950 FDTD stepper behind
 * photonfdtd's Simulation(backend="hip").
 *
 * The reference binds its GPU stepper through PyO3 (rust/src/lib.rs:139-152,
 * rust/src/cuda.rs:508-574: CudaStepper.new_f32/new_f64, run_steps,
 * read_field, device_info, CUDA_BUILT). This header is the plain-C
 * replacement: every entry point takes plain pointers and sizes, returns 0 on
 * success and a negative PF_E* code on failure, with a message available from
 * pf_last_error(). Handles are not thread-safe; one HIP stream per handle.
 *
 * Array conventions: fields are C-order (nx, ny, nz), z fastest. "dtype"
 * arrays are float (PF_F32) or double (PF_F64) according to pf_desc.dtype.
 */
#ifndef PFHIP_H
#define PFHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFHIP_ABI_VERSION 9

enum pf_dtype { PF_F32 = 0, PF_F64 = 1 };
enum pf_comp { PF_EX = 0, PF_EY = 1, PF_EZ = 2, PF_HX = 3, PF_HY = 4, PF_HZ = 5 };
enum pf_status {
  PF_OK = 0,
  PF_E_ARG = -1,      /* invalid argument / shape mismatch                 */
  PF_E_HIP = -2,      /* HIP runtime error (allocation, launch, copy)      */
  PF_E_STATE = -3,    /* call not valid in the handle's current state      */
  PF_E_LIMIT = -4     /* request exceeds a documented size limit           */
};

/* A strided sub-box of the grid: nodes start[a] + n*stride[a], n < count[a]. */
typedef struct pf_box {
  int64_t start[3];
  int64_t count[3];
  int64_t stride[3];
} pf_box;

/* One dispersive material (simulation.py:572-636): the cells it owns (flat
 * C-order indices, ascending) and, per pole, (a, b, c) of
 * P^{n+1} = a P^n + b P^{n-1} + c E^n. E is corrected by
 * -inv_eps * sum_p (P^{n+1} - P^n) (simulation.py:645-669). */
typedef struct pf_ade_desc {
  int64_t n_cells;
  const int64_t* cells;
  int32_t n_poles;
  const double* abc;      /* [n_poles][3] */
  double inv_eps;         /* 1 / (eps0 * eps_inf) */
} pf_ade_desc;

/* Running DFT over a box (simulation.py:1023-1042,1116-1133):
 * acc[c][f] += phase[s][f] * field_c * dt for every sampled step s, E
 * components with phase_e (t = (n+1) dt), H components with phase_h
 * (t = (n+1/2) dt). Steps n with n % interval == 0 are sampled, in order. */
typedef struct pf_dft_desc {
  int32_t n_comp;
  int32_t comps[6];
  pf_box box;
  int32_t n_freq;
  int32_t interval;
  int64_t n_samples;          /* number of sampled steps over the whole run */
  const double* phase_e;      /* [n_samples][n_freq][2] (re, im) */
  const double* phase_h;      /* [n_samples][n_freq][2] */
  double dt;                  /* Riemann weight */
} pf_dft_desc;

/* Time-domain Poynting flux through the plane index `index` normal to
 * `axis` (simulation.py:1440-1459): total += sum(S . n) * dA * dt per step. */
typedef struct pf_flux_desc {
  int32_t axis;
  int64_t index;
  double dA;
  double dt;
} pf_flux_desc;

/* Everything the stepper needs; host arrays are copied during pf_create.
 * Mirrors the tables rustbackend._build_tables builds (rustbackend.py:135-166)
 * plus the ADE, DFT and flux state the reference GPU path keeps on the host. */
typedef struct pf_desc {
  int32_t dtype;              /* pf_dtype */
  int32_t exact;              /* 1: the reference NumPy operation order
                                 (difference / cell size, no contraction);
                                 0: multiply by 1/cell size, fused kernels */
  int64_t shape[3];
  double cell[3];             /* cell size per axis (1.0 on collapsed axes) */
  double ch;                  /* dt / mu0 */
  const void* ce;             /* [nx*ny*nz] dt/(eps_r eps0), dtype */
  const double* b_e[3];       /* CPML coefficients per axis, [shape[a]] */
  const double* c_e[3];
  const double* b_h[3];
  const double* c_h[3];
  int64_t n_steps;            /* length of each waveform row */
  /* soft sources, sorted by (component, cell); equal targets are applied in
   * array order (the reference's source order) */
  int32_t n_src;
  const int32_t* src_comp;
  const int64_t* src_cell;
  const double* src_amp;
  const int32_t* src_wave;    /* row of `wave` holding w(t) at the Yee time */
  int32_t n_wave;
  const double* wave;         /* [n_wave][n_steps] */
  int32_t n_ade;
  const pf_ade_desc* ade;
  int32_t n_dft;
  const pf_dft_desc* dft;
  int32_t n_flux;
  const pf_flux_desc* flux;
  int32_t flags;              /* PF_FLAG_* */
  /* Subpixel smoothing (simulation.py:956-969): per-component E
   * coefficients dt/(eps_yy eps0) and dt/(eps_zz eps0) for Ey and Ez, in the
   * dtype; `ce` is then the Ex coefficient. NULL (both): ce for all three.
   * Such a handle steps with the generic two-pass kernels. */
  const void* ce_y;
  const void* ce_z;
  /* Beyond-HBM stepping (PF_FLAG_STREAM): steps per tile visit (0: 32) and
   * owned x planes per tile (0: sized from the free device memory). */
  int32_t t_block;
  int32_t tile_rows;
} pf_desc;

/* pf_desc.flags: use the in-place two-pass kernels only (no fused ping-pong
 * step); required for pf_checkpoint_enable / pf_adjoint. */
#define PF_FLAG_GENERIC 1
/* Keep the domain in pinned host memory and stream x-slab tiles through the
 * device, t_block steps per visit (the reference's StreamingCudaStepper,
 * rust/src/cuda.rs:576-1116, picked by rustbackend.py:264-290 when the run
 * does not fit the device). Same bits as the resident stepper. Supported:
 * CPML, soft sources, DFT monitors, field reads/writes; not: dispersive
 * media, flux monitors, subpixel coefficients, checkpoints / adjoint,
 * pf_field_ptr (the fields are not on the device). */
#define PF_FLAG_STREAM 2
/* Step with the 64-lane march on dense z rows (nz a multiple of 64, so a
 * field's layout is the generic kernels' C order): pf_checkpoint_enable /
 * pf_adjoint then work on the fused stepper, whose forward run is the march
 * (no reference counterpart; pf_adjoint's own steps are unchanged).
 * pf_checkpoint_enable refuses a handle that did not end up in that layout. */
#define PF_FLAG_DENSE64 4

/* Cotangent of one DFT accumulator: g = dJ/d(Re F) + i dJ/d(Im F) for monitor
 * `mon`, component slot `slot`, complex128 interleaved [n_freq][box cells]. */
typedef struct pf_cotangent {
  int32_t mon;
  int32_t slot;
  const double* g;
} pf_cotangent;

typedef struct pf_handle pf_handle;

/* Kernel classes reported by pf_timing_read. */
enum pf_kernel_class {
  PF_K_YEE_H = 0,      /* H half-step (curl E + CPML)                       */
  PF_K_YEE_E = 1,      /* E half-step (curl H + CPML)                       */
  PF_K_YEE_FUSED = 2,  /* one launch doing a whole H+E step                 */
  PF_K_AUX = 3,        /* per-step sources, ADE, flux and DFT sample capture */
  PF_K_YEE_PAIR = 4,   /* one launch doing two whole steps (temporal block) */
  PF_K_DFT_FOLD = 5,   /* FP64 DFT accumulation of the buffered samples     */
  PF_NKCLASS = 6
};

/* Replaces CudaStepper.new_f32/new_f64 (rust/src/cuda.rs:424-540). */
int pf_create(const pf_desc* desc, pf_handle** out);
/* Frees every device allocation of the handle. */
int pf_destroy(pf_handle* h);
/* Replaces CudaStepper.run_steps (rust/src/cuda.rs:298-341,543-548):
 * advances steps [step0, step0 + n), synchronous on return. */
int pf_advance(pf_handle* h, int64_t step0, int64_t n);
/* Replaces CudaStepper.read_field (rust/src/cuda.rs:354-356): copies one
 * component (comp 0..5 = Ex..Hz) into host memory of nx*ny*nz dtype. */
int pf_read_field(pf_handle* h, int32_t comp, void* dst);
/* Strided sub-box read for FieldMonitor downsample / plane_z
 * (simulation.py:1062-1078); dst holds count[0]*count[1]*count[2] dtype. */
int pf_read_box(pf_handle* h, int32_t comp, const pf_box* box, void* dst);
/* A strided sub-box quantised on the device for compressed FieldMonitor
 * storage; replaces the host quantisation of CompressedFieldSeries.append
 * (storage.py:128-141). *scale receives max |x| over the box (1 when that is
 * 0 or not finite); dst receives one value per node: bits 8 -> int8
 * rint(x / scale * 127), bits 16 -> float16 bits of x / scale, computed in
 * the field dtype (round half to even). */
int pf_read_box_q(pf_handle* h, int32_t comp, const pf_box* box, int32_t bits, void* dst,
                  double* scale);
/* Overwrites one component from host memory (nx*ny*nz dtype). */
int pf_write_field(pf_handle* h, int32_t comp, const void* src);
/* DFT accumulator of monitor `mon`, component slot `slot` (index into its
 * comps[]): complex128 interleaved [n_freq][box cells]. Replaces the host
 * accumulation of rustbackend.py:384-397. */
int pf_dft_read(pf_handle* h, int32_t mon, int32_t slot, double* dst);
/* Device-resident arrays for zero-copy handoff to torch or another device
 * library (the survey's pf_field_ptr; the reference hands fields and DFT
 * planes over through host arrays, rustbackend.py:297-304,384-397). ptr is a
 * device address, dtype PF_F32 / PF_F64, shape and strides (in elements,
 * row-major) fill the first ndim entries; stream is the handle's hipStream_t
 * (as an integer), on which the data was written and which has been
 * synchronized before the call returns. */
typedef struct pf_view {
  void* ptr;
  int32_t dtype;
  int32_t ndim;
  int64_t shape[6];
  int64_t strides[6];
  uint64_t stream;
} pf_view;
/* Field component comp (0..5: Ex Ey Ez Hx Hy Hz) of the current time level,
 * shape (nx, ny, nz), strides (ny*zrow, zrow, 1). Valid until the next
 * pf_advance, pf_write_field, pf_reset or pf_destroy on the handle; writes
 * through it change the simulation state. PF_E_STATE when the stepper keeps
 * its rows in padded blocks (62 or 60 nodes per 64 slots), which are not a
 * strided view. */
int pf_field_ptr(pf_handle* h, int32_t comp, pf_view* out);
/* DFT accumulators of monitor mon with every buffered sample folded in:
 * float64, shape (n_comp, n_freq, bx, by, bz, 2) (real, imaginary; components
 * in the order of the monitor's comps[]), contiguous. Valid until the next
 * pf_advance or pf_destroy. */
int pf_dft_ptr(pf_handle* h, int32_t mon, pf_view* out);
/* Folds the DFT samples the stepper still buffers into the complex128
 * accumulators of every monitor and waits for it: afterwards the
 * accumulators hold every step advanced so far (the reference accumulates
 * each sampled step at once, simulation.py:1116-1133; pf_dft_read folds
 * implicitly). */
int pf_dft_flush(pf_handle* h);
/* How the handle steps (no reference counterpart: the reference's CUDA
 * stepper has one kernel pair). path: 0 generic two-pass kernels, 1 fused
 * one-launch step on dense rows, 2 the 3-D march, 3 the march with the
 * two-step launch, 4 beyond-HBM streaming of x-slab tiles (tiles_bulk: the
 * number of tiles). zpad: nodes per 64-slot padded z block (0: dense rows;
 * 64: rows of the 64-lane march, node k in slot k), zrow: slots per row.
 * tiles_* / cells_*: march tiles (and the cells they own) that lie
 * off every CPML slab (bulk, the psi-free kernel) and that touch one (cpml).
 * cpml_lists: which CPML tile lists are non-empty, bit 0 the tiles in slabs
 * of several axes (or of any axis with padded rows), bits 1 / 2 / 3 the
 * tiles in z / y / x slabs only (dense rows: each runs the kernel variant of
 * that axis). dft_ring: samples a DFT monitor buffers before a fold (largest
 * over monitors). */
typedef struct pf_info {
  int32_t path;
  int32_t zpad;
  int32_t zrow;
  int32_t march_nw;
  int32_t pml_nw;
  int32_t cpml_lists;
  int64_t tiles_bulk;
  int64_t tiles_cpml;
  int64_t cells_bulk;
  int64_t cells_cpml;
  int64_t pair_tiles_bulk;
  int64_t pair_tiles_cpml;
  int64_t dft_ring;
  int64_t post_fused;   /* 1: ADE, E sources and DFT samples in one launch per step */
} pf_info;
int pf_stepper_info(pf_handle* h, pf_info* out);
/* Accumulated flux of every flux monitor, [n_flux] doubles. */
int pf_flux_read(pf_handle* h, double* dst);
/* Discrete adjoint. The reference differentiates its stepper through JAX
 * (jaxbackend.py:666-722, value_and_grad_eps); these two calls are the native
 * replacement. pf_checkpoint_enable(h, C) must precede the forward run from
 * step 0: pf_advance then stores the full state every C steps on the device.
 * After the forward run reaches the end of the waveform table, pf_adjoint
 * sweeps backwards (recomputing each C-step segment from its checkpoint) with
 * the DFT cotangents as time-reversed sources and writes dJ/dce for every
 * node (float64, nx*ny*nz) to grad_ce. The cotangent arrays and grad_ce may
 * be host or device memory. Dispersive media are rejected. */
int pf_checkpoint_enable(pf_handle* h, int32_t interval);
int pf_adjoint(pf_handle* h, int32_t n_cot, const pf_cotangent* cot, double* grad_ce);
/* Batched 2-D stepping (no reference counterpart: the reference runs one
 * simulation per Simulation.run). A batch holds n independent xy-plane
 * simulations (nz == 1) of one dtype, each described by a pf_desc, with
 * pols[i] = 0 (TE: Ex, Ey, Hz) or 1 (TM: Ez, Hx, Hy) the group its sources
 * excite. Supported per simulation: CPML, soft sources, ADE media and DFT
 * monitors (no flux). pf_batch_advance(b, n) advances every simulation n
 * steps (a simulation stops at its own step count). One march kernel does
 * a whole time step; each wavefront owns a 62-column strip over `rows` rows.
 * Modes: 0 one march launch per step, 1 one persistent launch for the whole
 * advance (grid barrier with a bounded wait: PF_E_STATE if the workgroups
 * were not all resident), 2 march launches replayed from a captured
 * hipGraph. waves_per_block is 1, 2, 4 or 8. pf_batch_reset zeroes every
 * field, psi, ADE and DFT state and the step counter (the autotuner times
 * candidate settings from a reset state). pf_batch_persistent_grid reports
 * the occupancy and the workgroup count mode 1 would use, without
 * launching.
 * A batch of 3-D descriptors (every shape[2] >= 2; pols ignored) runs, per
 * step, the single-run march kernels over the tiles of all simulations
 * still running (one psi-free and one CPML launch), then one ADE, one
 * E-source and one DFT-capture launch for all of them: mode 0 only, waves_per_block 10, 12 or 16 (march
 * wavefronts per workgroup), rows 1..128 (x planes per psi-free tile). Each 3-D
 * simulation needs the march kernel's conditions (E-side sources only,
 * CPML slabs) and no flux planes. 2-D and 3-D descriptors do not mix. */
typedef struct pf_batch pf_batch;
int pf_batch_create(const pf_desc* descs, const int32_t* pols, int32_t n, pf_batch** out);
int pf_batch_destroy(pf_batch* b);
int pf_batch_configure(pf_batch* b, int32_t mode, int32_t waves_per_block, int32_t rows);
int pf_batch_reset(pf_batch* b);
int pf_batch_persistent_grid(pf_batch* b, int32_t waves_per_block, int32_t* per_cu,
                             int32_t* n_cu, int32_t* grid);
int pf_batch_advance(pf_batch* b, int64_t n);
/* Device time (HIP events on the batch stream) of the last pf_batch_advance. */
int pf_batch_last_ms(pf_batch* b, double* ms);
/* Replace ce = dt / (eps_r eps0) of simulation `sim` (nx*ny values of the
 * batch dtype): a new permittivity for the next run after pf_batch_reset. */
int pf_batch_write_ce(pf_batch* b, int32_t sim, const void* ce);
int pf_batch_dft_read(pf_batch* b, int32_t sim, int32_t mon, int32_t slot, double* dst);
int pf_batch_read_field(pf_batch* b, int32_t sim, int32_t comp, void* dst);
/* Event timing (no reference counterpart; used by the benchmark): when on,
 * each kernel-class launch group of pf_advance is bracketed by HIP events on
 * the handle's stream. on = 0: off; 1: every group; k > 1: the march and aux
 * groups of every k-th single step only (their totals and counts cover those
 * steps; the other classes are timed in every step), which keeps the event
 * packets out of most steps. Enabling resets the totals. pf_timing_read
 * returns, per class, total milliseconds and launch counts ([PF_NKCLASS]
 * each). */
int pf_timing_enable(pf_handle* h, int32_t on);
int pf_timing_read(pf_handle* h, double* ms, int64_t* launches);
/* Replaces CudaStepper.device_info (rust/src/cuda.rs:560-573). */
int pf_device_info(char* name, int32_t name_len, int64_t* free_mb, int64_t* total_mb);
/* Thread-local description of the last failure on this thread. */
const char* pf_last_error(void);
/* PFHIP_ABI_VERSION of the loaded library (the CUDA_BUILT analogue). */
int32_t pf_abi_version(void);

/* ---- Learned surrogate kernels (no reference counterpart) -------------
 * Layer kernels of the spectral surrogate on the matrix cores (BF16
 * operands, FP32 accumulation). Tensors are device pointers in NHWC order:
 * activations BF16 (raw 16-bit), gradients FP32, weights BF16 [cout][k][k][cin]
 * with FP32 master copies kept by the optimizer. `stream` is a hipStream_t
 * (NULL: the default stream). Calls are asynchronous on that stream and
 * return PF_E_ARG for bad shapes, PF_E_HIP for launch errors
 * (pf_nn_last_error gives the text). */
typedef struct pf_nn_conv {
  int32_t n, h, w;       /* input batch and spatial size */
  int32_t cin, cout;
  int32_t ksize;         /* 1, or 3 with padding 1 */
  int32_t stride;        /* 1 or 2; output size (h - 1) / stride + 1 */
  int32_t relu;          /* 1: ReLU after the bias */
  int32_t out_f32;       /* 1: output stored as FP32, else BF16 */
} pf_nn_conv;
int pf_nn_conv_forward(const pf_nn_conv* c, const void* x, const void* w, const float* bias,
                       void* y, void* stream);
/* dy: FP32 gradient w.r.t. the layer output y; dx (FP32, may be NULL) is
 * written; dw and dbias (FP32, may be NULL) are accumulated into. The weight
 * gradient sums fixed pixel splits in a fixed order (same bits every run);
 * its partial sums use a workspace allocated and freed in order on the
 * stream (hipMallocAsync / hipFreeAsync). */
int pf_nn_conv_backward(const pf_nn_conv* c, const void* x, const void* w, const void* y,
                        const float* dy, float* dx, float* dw, float* dbias, void* stream);
/* pf_nn_conv_backward without any allocation (no reference counterpart:
 * the reference has no learned component). For callers that cannot
 * allocate device memory: code under HIP graph capture, tight training or
 * serving loops. Same arguments, kernels and bits as pf_nn_conv_backward;
 * the weight gradient's per-split partial sums go to a static 16 MiB device
 * array per GPU instead of a stream-ordered workspace. PF_E_LIMIT when a
 * layer needs more (cout * ksize^2 * cin above 4 Mi floats). Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams,
 * other threads) race. The stream must belong to the current device. */
int pf_nn_conv_backward_noalloc(const pf_nn_conv* c, const void* x, const void* w, const void* y,
                                const float* dy, float* dx, float* dw, float* dbias, void* stream);
/* loss (one FP32, zeroed by the caller) += mean (pred - target)^2; dpred written. */
int pf_nn_mse(int32_t total, const float* pred, const float* target, float* dpred, float* loss,
              void* stream);
/* Adam step `step` (1-based) on FP32 master weights; refreshes the BF16 copy
 * (may be NULL) and zeroes grad. */
int pf_nn_adam(int64_t n, float* param, float* grad, float* m, float* v, void* param_bf16,
               float lr, float beta1, float beta2, float eps, int32_t step, void* stream);
int pf_nn_to_bf16(int64_t n, const float* x, void* y, void* stream);
const char* pf_nn_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PFHIP_H */
