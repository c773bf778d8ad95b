// sph_device.h -- device-side SPHENIX physics for gfx950.
//
// Constants restate src/kernel_hydro.h:121-146 and 203-241 (Wendland C2, 3D),
// src/dimension.h:40-41, src/adiabatic_index.h:41-49 (gamma = 5/3) and
// src/hydro/SPHENIX/hydro_parameters.h:53 (beta_visc = 3).  The constant
// expressions are evaluated exactly as the reference's macros are (float
// products, double reciprocals rounded to float) so that kernel_gamma2,
// kernel_root and the pair test agree bit-for-bit with the host.
//
// The per-pair bodies compute the same quantities as
// src/hydro/SPHENIX/hydro_iact.h (runner_iact_nonsym_density :141,
// runner_iact_nonsym_gradient :291, runner_iact_nonsym_force :507) but are
// arranged for the VALU: the Wendland polynomial is evaluated in its factored
// form (1-q)^4 (1+4q), one reciprocal square root gives both r and 1/r, and
// every per-particle quotient (P/rho^2, 1/rho, 1/m, h^-4) is hoisted out of
// the pair loop.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

// --- kernel_hydro.h constants (evaluated as the reference's macros are) ---
constexpr float kGamma = (float)(1.936492);
constexpr float kConst = (float)(21. * 0.31830988618379067154 / 2.);  // 21/(2 pi)
constexpr float kGammaInv = (float)(1. / (double)kGamma);
constexpr float kGamma2 = kGamma * kGamma;
constexpr float kGammaInvDim = (float)(1. / (double)(kGamma * kGamma * kGamma));
constexpr float kGammaInvDim1 =
    (float)(1. / (double)(kGamma * kGamma * kGamma * kGamma));
constexpr float kRoot = 1.f * kConst * kGammaInvDim;
// Factors folded into the factored polynomial: W = cW (1-q)^4 (1+4q),
// dW/du = cdW q (1-q)^3 with q = u / gamma.
constexpr float cW = kConst * kGammaInvDim;
constexpr float cdW = -20.f * kConst * kGammaInvDim1;

constexpr float kDim = 3.f;
constexpr float kDimInv = 0.3333333333f;
constexpr float kGasGamma = 1.66666666666666667f;
constexpr float kGasGm1 = 0.66666666666666667f;
constexpr float kViscBeta = 3.0f;
constexpr int kNumTimeBins = 56;
constexpr int kBinInhibited = kNumTimeBins + 2;

// Wendland C2 W(u) and dW/du for u = r/h (kernel_deval, kernel_hydro.h:257),
// factored form; q >= 1 gives zero.
__device__ __forceinline__ void kernel_deval(float u, float &w, float &dw) {
  const float q = u * kGammaInv;
  const float om = fmaxf(1.f - q, 0.f);
  const float om2 = om * om;
  const float om3 = om2 * om;
  w = (om2 * om2) * fmaf(4.f * cW, q, cW);
  dw = (q * om3) * cdW;
}

// Only the derivative (gradient loop).
__device__ __forceinline__ float kernel_dw(float u) {
  const float q = u * kGammaInv;
  const float om = fmaxf(1.f - q, 0.f);
  return (q * (om * om * om)) * cdW;
}

// Frame of a pair of adjacent top-level cells, as space_getsid_and_swap_cells
// finds it (space_getsid.h:47-80): the sort id of the cells' offset and the
// swap (runner_flip, sort_part.h:59-61) that makes the second cell of the
// pair the one in the positive half-space of the offset.  The pair loops
// take both cells' particles relative to that upper cell's corner, with the
// periodic shift on the other cell's side (runner_doiact_functions_hydro.h:
// 1327-1341, 1672-1674; tools.c:218-220).  For cells adjacent on a grid of
// >= 3 cells per axis the wrapped loc difference of getsid is the stencil
// offset o times the width and its shift is the image translation t of j's
// cell, so, with loc = index * width (space_regrid.c:307):
//   sort id 13 (o = 0)           the same cell: a self pair;
//   sort id > 13 (no swap)       x_i relative to oi = loc_j + t,
//                                x_j relative to oj = loc_j;
//   sort id < 13 (swap)          x_i relative to oi = loc_i,
//                                x_j relative to oj = loc_i + (-t).
// Writes fo = {oi[3], oj[3]}; returns true for a self pair.
__device__ __forceinline__ bool pair_frame(const int o[3], const int ci[3],
                                           const int cj[3], const double t[3],
                                           const double w[3], double *fo) {
  const int sid = 9 * (o[0] + 1) + 3 * (o[1] + 1) + (o[2] + 1);
  const bool up_j = sid > 13;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double li = (double)ci[k] * w[k], lj = (double)cj[k] * w[k];
    fo[k] = up_j ? lj + t[k] : li;
    fo[3 + k] = up_j ? lj : li + (-t[k]);
  }
  return sid == 13;
}

// The FP32 pair contract, the expression of the reference's loops and
// brute-force checkers: in one cell (self) the FP64 difference rounded once
// (DOSELF1 :2459, DOSELF2 :2784, tools.c:617-619); across cells both
// cell-frame offsets rounded to float and subtracted in float (DOPAIR1
// :1327-1346, DOPAIR2 :1748-1781, tools.c:232-246); r2 as two FMAs, the
// sum of squares as gcc contracts it.  fo = {oi[3], oj[3]} of pair_frame.
__device__ __forceinline__ float pair_r2(bool self, double xi, double yi,
                                         double zi, double xj, double yj,
                                         double zj, const double *fo,
                                         float &dx, float &dy, float &dz) {
  if (self) {
    dx = (float)(xi - xj);
    dy = (float)(yi - yj);
    dz = (float)(zi - zj);
  } else {
    dx = (float)(xi - fo[0]) - (float)(xj - fo[3]);
    dy = (float)(yi - fo[1]) - (float)(yj - fo[4]);
    dz = (float)(zi - fo[2]) - (float)(zj - fo[5]);
  }
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

__device__ __forceinline__ float search_radius2(float h) {
  return (h * h) * kGamma2;
}

// Radius of the FP32 pre-test for an exact squared radius R.  The pre-test
// distance is formed from float positions relative to the candidate's
// top-level cell corner; each component differs from the pair contract's
// by at most eps (both are within a few float roundings of the real
// difference, and all terms are bounded by a few cell widths).  Then
// |r_f^2 - r^2| <= eps (2 sqrt(3) r + 3 eps) + 2 x 2^-22 r^2 (rounding of
// both squares), so r^2 < R implies r_f^2 < pretest_radius2(R, eps).
__device__ __forceinline__ float pretest_radius2(float R, float eps) {
  return fmaf(eps, fmaf(4.f, sqrtf(R), 4.f * eps), R * 1.000002f);
}

__device__ __forceinline__ float fast_rsq(float x) {
  return __builtin_amdgcn_rsqf(x);
}
__device__ __forceinline__ float fast_rcp(float x) {
  return __builtin_amdgcn_rcpf(x);
}
__device__ __forceinline__ float fast_sqrt(float x) {
  return __builtin_amdgcn_sqrtf(x);
}

// Per-particle kernels over [lo, hi): a block covers kRuns consecutive runs
// of blockDim.x particles (launch run_blocks(hi - lo, blockDim.x) blocks).
// The keys (time bin, redo flag) of all runs are loaded before any particle
// is processed, so every thread has kRuns loads in flight: with one byte
// load per thread a sweep over every particle slot is latency-bound, which
// is most of the cost of a step with few active particles.
constexpr int kRuns = 8;
__host__ __device__ inline int run_blocks(int n, int block) {
  return n > 0 ? (n + block * kRuns - 1) / (block * kRuns) : 0;
}
template <class K, class F>
__device__ __forceinline__ void for_runs(int lo, int hi,
                                         const K *__restrict__ key, F f) {
  const int base = lo + blockIdx.x * blockDim.x * kRuns + threadIdx.x;
  K k[kRuns];
#pragma unroll
  for (int u = 0; u < kRuns; u++) {
    const int i = base + u * (int)blockDim.x;
    k[u] = i < hi ? key[i] : K{};
  }
#pragma unroll
  for (int u = 0; u < kRuns; u++) {
    const int i = base + u * (int)blockDim.x;
    if (i < hi) f(i, k[u]);
  }
}

// The same over the entries of a compact particle list (list[0 .. *count),
// e.g. the work queue's list of the particles a loop processed), launched
// with run_blocks(capacity, blockDim.x) blocks: entries past the count are
// skipped, so a list much shorter than the particle slots costs little.
template <class K, class F>
__device__ __forceinline__ void for_runs_list(const int *__restrict__ list,
                                              const int *__restrict__ count,
                                              const K *__restrict__ key, F f) {
  const int n = *count;
  const int base = blockIdx.x * blockDim.x * kRuns + threadIdx.x;
  if (base >= n) return;
  int idx[kRuns];
#pragma unroll
  for (int u = 0; u < kRuns; u++) {
    const int k = base + u * (int)blockDim.x;
    idx[u] = k < n ? list[k] : -1;
  }
  K k[kRuns];
#pragma unroll
  for (int u = 0; u < kRuns; u++) k[u] = idx[u] >= 0 ? key[idx[u]] : K{};
#pragma unroll
  for (int u = 0; u < kRuns; u++)
    if (idx[u] >= 0) f(idx[u], k[u]);
}

}  // namespace sph
