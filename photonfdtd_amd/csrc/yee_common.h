// Device helpers shared by the generic and batched per-node Yee kernels.
#pragma once

#include "pfhip_internal.h"

namespace pfhip {

// Flat psi index of a compact PML array whose axis `ax` holds only the
// `pc` nodes with a non-zero CPML c coefficient (rust/src/kernel.rs:92-103).
__device__ __forceinline__ int64_t psi_index(int ax, int m, int i, int j, int k, int ny,
                                             int nz, int pc) {
  if (ax == 0) return ((int64_t)m * ny + j) * nz + k;
  if (ax == 1) return ((int64_t)i * pc + m) * nz + k;
  return ((int64_t)i * ny + j) * pc + m;
}

template <typename real, bool EXACT>
__device__ __forceinline__ real diff(real hi, real lo, real d, real inv_d) {
  if constexpr (EXACT) return (hi - lo) / d;
  else return (hi - lo) * inv_d;
}

// One derivative term of a curl: present flag, value, and PML state.
template <typename real>
struct Term {
  bool on;
  real d;
  real* psi;    // nullptr when the node is outside the PML slab
  int64_t pid;
  real b, c;
};

// F += K * (dA - dB), then F += K psiA, F -= K psiB on PML nodes, where
// psi <- b psi + c d first. Same order as the reference's bulk-then-slab
// update (simulation.py:1263-1309).
template <typename real>
__device__ __forceinline__ void apply_terms(real* __restrict__ F, int64_t id, real K,
                                            const Term<real>& A, const Term<real>& B) {
  if (!A.on && !B.on) return;
  real cb;
  if (A.on) {
    cb = A.d;
    if (B.on) cb = cb - B.d;
  } else {
    cb = -B.d;
  }
  real f = F[id];
  f = f + cb * K;
  if (A.on && A.psi) {
    real p = A.b * A.psi[A.pid] + A.c * A.d;
    A.psi[A.pid] = p;
    f = f + K * p;
  }
  if (B.on && B.psi) {
    real p = B.b * B.psi[B.pid] + B.c * B.d;
    B.psi[B.pid] = p;
    f = f - K * p;
  }
  F[id] = f;
}

// One curl component: f + K (dA - dB) with the PML slab terms after the bulk
// term (same order as the generic kernels).
template <typename real>
__device__ __forceinline__ real curl_update(real f, real K, bool onA, real dA, bool pmlA,
                                            real psiA, real bA, real cA, real* newA, bool onB,
                                            real dB, bool pmlB, real psiB, real bB, real cB,
                                            real* newB) {
  if (!onA && !onB) return f;
  real cb;
  if (onA) {
    cb = dA;
    if (onB) cb = cb - dB;
  } else {
    cb = -dB;
  }
  f = f + cb * K;
  if (onA && pmlA) {
    const real p = bA * psiA + cA * dA;
    *newA = p;
    f = f + K * p;
  }
  if (onB && pmlB) {
    const real p = bB * psiB + cB * dB;
    *newB = p;
    f = f - K * p;
  }
  return f;
}

}  // namespace pfhip
