// sph_loops.h -- the three neighbour-loop policies (density, gradient,
// force): per-lane particle state, accumulators, the pair test and the
// interaction body, each restating one runner_iact_nonsym_* function.
#pragma once

#include "sph_device.h"
#include "sph_kernels.h"

// Occupancy targets (waves per SIMD) of the batched neighbour kernels.  The
// one-tile-per-wave kernel needs fewer registers than the grid and
// persistent forms: its force loop fits 6 waves per SIMD without spilling
// (the others spill 16-31 VGPRs there).
constexpr int kOccDensity = 7;
constexpr int kOccGradient = 7;
constexpr int kOccForce = 5;
constexpr int kOccForceTiles = 6;
// The sparse-tile kernels (sph_sparse.h): one tile per wave.
constexpr int kOccSparse = 6;  // density, gradient (their LDS allows 6)
constexpr int kOccForceSparse = 5;
// The persistent forms of both engines: their loop around the tile costs
// registers (at 7 waves per SIMD the density form spilled 34 VGPRs).
constexpr int kOccSparseQueue = 5;

namespace sphl {
using namespace sph;

__device__ __forceinline__ bool is_active(int8_t tb, int max_active_bin) {
  return tb <= max_active_bin && tb != kBinInhibited;  // active.h:349, 518
}

// Position of an inhibited particle in its records: every test rejects it.
constexpr double kFar = 1.0e30;

// ------------------------------------------------------ neighbour records --
// The fields a neighbour loop reads from particle j, from the particle
// arrays (rebuilt after the stage that changes them).
__device__ __forceinline__ DensRec dens_rec(const DevParts &d, int i) {
  const bool inh = d.tb[i] == kBinInhibited;
  DensRec r;
  r.x = inh ? kFar : d.px[i];
  r.y = inh ? kFar : d.py[i];
  r.z = inh ? kFar : d.pz[i];
  r.vx = d.vx[i];
  r.vy = d.vy[i];
  r.vz = d.vz[i];
  r.m = d.m[i];
  return r;
}

// The gradient record of a particle (not inhibited) whose rho, u, c and
// alpha the caller holds (the ghost, as it finishes the particle).
__device__ __forceinline__ GradRec grad_rec_of(const DevParts &d, int i,
                                               float rho, float u, float c,
                                               float alpha) {
  GradRec r;
  r.x = d.px[i];
  r.y = d.py[i];
  r.z = d.pz[i];
  r.vx = d.vx[i];
  r.vy = d.vy[i];
  r.vz = d.vz[i];
  r.mrho = rho > 0.f ? d.m[i] / rho : 0.f;
  r.u = u;
  r.c = c;
  r.alpha = alpha;
  r.pad0 = 0.f;
  return r;
}

__device__ __forceinline__ GradRec grad_rec(const DevParts &d, int i) {
  GradRec r = grad_rec_of(d, i, d.rho[i], d.u[i], d.c[i], d.alpha[i]);
  if (d.tb[i] == kBinInhibited) r.x = r.y = r.z = kFar;
  return r;
}

// The force record of a particle (not inhibited) whose h, rho, P, c,
// alpha, alpha_diff and u the caller holds (the extra ghost).
__device__ __forceinline__ ForceRec force_rec_of(const DevParts &d, int i,
                                                 int8_t tb, float h, float rho,
                                                 float P, float c, float alpha,
                                                 float adiff, float u) {
  ForceRec r;
  r.x = d.px[i];
  r.y = d.py[i];
  r.z = d.pz[i];
  r.hg2 = search_radius2(h);
  r.vx = d.vx[i];
  r.vy = d.vy[i];
  r.vz = d.vz[i];
  r.m = d.m[i];
  r.inv_m = 1.f / r.m;
  r.h_inv = 1.f / h;
  r.rho = rho;
  r.inv_rho = 1.f / rho;
  r.P = P;
  r.pr2 = P / (rho * rho);
  r.f = d.f[i];
  r.bal = d.bal[i];
  r.c = c;
  r.alpha = alpha;
  r.Pad = P * adiff;
  r.u = u;
  r.tb = (float)tb;
  return r;
}

__device__ __forceinline__ ForceRec force_rec(const DevParts &d, int i) {
  const int8_t tb = d.tb[i];
  ForceRec r = force_rec_of(d, i, tb, d.h[i], d.rho[i], d.P[i], d.c[i],
                            d.alpha[i], d.adiff[i], d.u[i]);
  if (tb == kBinInhibited) {
    r.x = r.y = r.z = kFar;
    r.hg2 = -1.f;
  }
  return r;
}

// Candidate record of the FP32 pre-test: position relative to the corner of
// the particle's cell and the widened search radius.
__device__ __forceinline__ PosRec pos_rec(const DevParts &d,
                                          const GridDesc &g, float pos_eps,
                                          int i) {
  const bool inh = d.tb[i] == kBinInhibited;
  const int c = d.cell[i];
  const int cz = c % g.nc[2], cy = (c / g.nc[2]) % g.nc[1];
  const int cx = c / (g.nc[2] * g.nc[1]);
  PosRec p;
  p.x = inh ? 1e30f : (float)(d.px[i] - (double)cx * (g.dim[0] / g.nc[0]));
  p.y = inh ? 1e30f : (float)(d.py[i] - (double)cy * (g.dim[1] / g.nc[1]));
  p.z = inh ? 1e30f : (float)(d.pz[i] - (double)cz * (g.dim[2] / g.nc[2]));
  p.hg2m = inh ? -1.f : pretest_radius2(search_radius2(d.h[i]), pos_eps);
  return p;
}

// ---------------------------------------------------------------- density --
struct DensAcc {
  float rho, rho_dh, wc, wc_dh, divv, rx, ry, rz;
  int n;
};
struct DensI {
  double x, y, z;
  float vx, vy, vz, h_inv, hg2;
};

struct DensityLoop {
  // occupancy target of the batched kernels (waves per SIMD)
  static constexpr int kWavesPerSimd = kOccDensity;
  static constexpr int kTileWavesPerSimd = kOccDensity;
  static constexpr int kSparseWavesPerSimd = kOccSparse;
  using Rec = DensRec;
  using Acc = DensAcc;
  using I = DensI;
  static __device__ void load_i(const DevParts &d, int i, I &s) {
    s.x = d.px[i];
    s.y = d.py[i];
    s.z = d.pz[i];
    s.vx = d.vx[i];
    s.vy = d.vy[i];
    s.vz = d.vz[i];
    const float h = d.h[i];
    s.h_inv = 1.f / h;
    s.hg2 = search_radius2(h);
  }
  static __device__ void load_acc(const DevParts &d, int i, Acc &a) {
    a.rho = d.rho[i];
    a.rho_dh = d.rho_dh[i];
    a.wc = d.wc[i];
    a.wc_dh = d.wc_dh[i];
    a.divv = d.divv[i];
    a.rx = d.rotx[i];
    a.ry = d.roty[i];
    a.rz = d.rotz[i];
    a.n = d.nd[i];
  }
  static __device__ void store_acc(const DevParts &d, int i, const Acc &a) {
    d.rho[i] = a.rho;
    d.rho_dh[i] = a.rho_dh;
    d.wc[i] = a.wc;
    d.wc_dh[i] = a.wc_dh;
    d.divv[i] = a.divv;
    d.rotx[i] = a.rx;
    d.roty[i] = a.ry;
    d.rotz[i] = a.rz;
    d.nd[i] = a.n;
  }
  static constexpr bool kSymmetric = false;  // search radius is H_i only
  static __device__ __forceinline__ void identity(Acc &a) {
    a.rho = a.rho_dh = a.wc = a.wc_dh = a.divv = a.rx = a.ry = a.rz = 0.f;
    a.n = 0;
  }
  // hydro_init_part's zero sums (hydro.h:564-585)
  static __device__ __forceinline__ void fresh_acc(const I &, Acc &a) {
    identity(a);
  }
  // Sum the partial sums of the S streams of a particle.
  template <int S>
  static __device__ __forceinline__ void combine(Acc &a) {
#pragma unroll
    for (int off = 64 / S; off < 64; off <<= 1) {
      a.rho += __shfl_xor(a.rho, off);
      a.rho_dh += __shfl_xor(a.rho_dh, off);
      a.wc += __shfl_xor(a.wc, off);
      a.wc_dh += __shfl_xor(a.wc_dh, off);
      a.divv += __shfl_xor(a.divv, off);
      a.rx += __shfl_xor(a.rx, off);
      a.ry += __shfl_xor(a.ry, off);
      a.rz += __shfl_xor(a.rz, off);
      a.n += __shfl_xor(a.n, off);
    }
  }
  // Candidate test: r < H_i (DOPAIR1 :1386, DOSELF1 :2396).
  template <class R>
  static __device__ __forceinline__ bool test(float r2, const I &s, const R *) {
    return r2 < s.hg2;
  }
  // runner_iact_nonsym_density (hydro_iact.h:141-193)
  static __device__ __forceinline__ void interact(float r2, float dx, float dy,
                                                  float dz, const I &s,
                                                  const Rec &j, Acc &a) {
    const float r_inv = r2 > 0.f ? fast_rsq(r2) : 0.f;
    const float r = r2 * r_inv;
    const float ui = r * s.h_inv;
    float wi, wi_dx;
    kernel_deval(ui, wi, wi_dx);
    const float t = fmaf(kDim, wi, ui * wi_dx);
    a.rho = fmaf(j.m, wi, a.rho);
    a.rho_dh = fmaf(-j.m, t, a.rho_dh);
    a.wc += wi;
    a.wc_dh -= t;
    const float faci = j.m * wi_dx * r_inv;
    const float dvx = s.vx - j.vx, dvy = s.vy - j.vy, dvz = s.vz - j.vz;
    const float dvdr = fmaf(dvz, dz, fmaf(dvy, dy, dvx * dx));
    a.divv = fmaf(-faci, dvdr, a.divv);
    a.rx = fmaf(faci, fmaf(dvy, dz, -dvz * dy), a.rx);
    a.ry = fmaf(faci, fmaf(dvz, dx, -dvx * dz), a.ry);
    a.rz = fmaf(faci, fmaf(dvx, dy, -dvy * dx), a.rz);
    a.n += 1;
  }
};

// --------------------------------------------------------------- gradient --
struct GradAcc {
  float vsig, lapu, avmax;
  int n;
};
struct GradI {
  double x, y, z;
  float vx, vy, vz, h_inv, hg2, u, c, alpha;
};

struct GradientLoop {
  // occupancy target of the batched kernels (waves per SIMD)
  static constexpr int kWavesPerSimd = kOccGradient;
  static constexpr int kTileWavesPerSimd = kOccGradient;
  static constexpr int kSparseWavesPerSimd = kOccSparse;
  using Rec = GradRec;
  using Acc = GradAcc;
  using I = GradI;
  static __device__ void load_i(const DevParts &d, int i, I &s) {
    s.x = d.px[i];
    s.y = d.py[i];
    s.z = d.pz[i];
    s.vx = d.vx[i];
    s.vy = d.vy[i];
    s.vz = d.vz[i];
    const float h = d.h[i];
    s.h_inv = 1.f / h;
    s.hg2 = search_radius2(h);
    s.u = d.u[i];
    s.c = d.c[i];
    s.alpha = d.alpha[i];
  }
  // The sums of a particle whose stencil one wave covers: hydro_prepare_-
  // gradient's initial v_sig = 2 c and alpha_visc_max_ngb = alpha
  // (hydro.h:671-750), the rest zero.
  static __device__ __forceinline__ void fresh_acc(const I &s, Acc &a) {
    a.vsig = 2.f * s.c;
    a.lapu = 0.f;
    a.avmax = s.alpha;
    a.n = 0;
  }
  static __device__ void load_acc(const DevParts &d, int i, Acc &a) {
    a.vsig = d.vsig[i];
    a.lapu = d.lapu[i];
    a.avmax = d.avmax[i];
    a.n = d.ng[i];
  }
  static __device__ void store_acc(const DevParts &d, int i, const Acc &a) {
    d.vsig[i] = a.vsig;
    d.lapu[i] = a.lapu;
    d.avmax[i] = a.avmax;
    d.ng[i] = a.n;
  }
  static constexpr bool kSymmetric = false;
  // v_sig >= 0 and alpha >= 0, so 0 is the identity of both maxima.
  static __device__ __forceinline__ void identity(Acc &a) {
    a.vsig = 0.f;
    a.lapu = 0.f;
    a.avmax = 0.f;
    a.n = 0;
  }
  template <int S>
  static __device__ __forceinline__ void combine(Acc &a) {
#pragma unroll
    for (int off = 64 / S; off < 64; off <<= 1) {
      a.vsig = fmaxf(a.vsig, __shfl_xor(a.vsig, off));
      a.lapu += __shfl_xor(a.lapu, off);
      a.avmax = fmaxf(a.avmax, __shfl_xor(a.avmax, off));
      a.n += __shfl_xor(a.n, off);
    }
  }
  template <class R>
  static __device__ __forceinline__ bool test(float r2, const I &s, const R *) {
    return r2 < s.hg2;
  }
  // runner_iact_nonsym_gradient (hydro_iact.h:291-344), a = 1, H = 0
  static __device__ __forceinline__ void interact(float r2, float dx, float dy,
                                                  float dz, const I &s,
                                                  const Rec &j, Acc &a) {
    const float r_inv = r2 > 0.f ? fast_rsq(r2) : 0.f;
    const float r = r2 * r_inv;
    const float dvdr =
        fmaf(s.vz - j.vz, dz, fmaf(s.vy - j.vy, dy, (s.vx - j.vx) * dx));
    const float mu = r_inv * fminf(dvdr, 0.f);
    a.vsig = fmaxf(a.vsig, fmaf(-kViscBeta, mu, s.c + j.c));
    const float dw = kernel_dw(r * s.h_inv);
    a.lapu = fmaf(j.mrho * (s.u - j.u) * r_inv, dw, a.lapu);
    a.avmax = fmaxf(a.avmax, j.alpha);
    a.n += 1;
  }
};

// ------------------------------------------------------------------ force --
struct ForceAcc {
  float ax, ay, az, udt, hdt;
  int n, mintb;
};
struct ForceI {
  double x, y, z;
  float vx, vy, vz, inv_m, hg2, h_inv, h_inv4;
  float inv_rho, rho, pr2, f, bal, c, alpha, P, Pad, u;
};

struct ForceLoop {
  // occupancy target of the batched kernels (waves per SIMD)
  static constexpr int kWavesPerSimd = kOccForce;
  static constexpr int kTileWavesPerSimd = kOccForceTiles;
  static constexpr int kSparseWavesPerSimd = kOccForceSparse;
  using Rec = ForceRec;
  using Acc = ForceAcc;
  using I = ForceI;
  static __device__ void load_i(const DevParts &d, int i, I &s) {
    const ForceRec &r = d.rf[rec_slot(d, i)];
    s.x = r.x;
    s.y = r.y;
    s.z = r.z;
    s.vx = r.vx;
    s.vy = r.vy;
    s.vz = r.vz;
    s.inv_m = r.inv_m;
    s.hg2 = r.hg2;
    s.h_inv = r.h_inv;
    s.h_inv4 = (r.h_inv * r.h_inv) * (r.h_inv * r.h_inv);
    s.inv_rho = r.inv_rho;
    s.rho = r.rho;
    s.pr2 = r.pr2;
    s.f = r.f;
    s.bal = r.bal;
    s.c = r.c;
    s.alpha = r.alpha;
    s.P = r.P;
    s.Pad = r.Pad;
    s.u = r.u;
  }
  static __device__ void load_acc(const DevParts &d, int i, Acc &a) {
    a.ax = d.ax[i];
    a.ay = d.ay[i];
    a.az = d.az[i];
    a.udt = d.udt[i];
    a.hdt = d.hdt[i];
    a.n = d.nf[i];
    a.mintb = d.mintb[i];
  }
  static __device__ void store_acc(const DevParts &d, int i, const Acc &a) {
    d.ax[i] = a.ax;
    d.ay[i] = a.ay;
    d.az[i] = a.az;
    d.udt[i] = a.udt;
    d.hdt[i] = a.hdt;
    d.nf[i] = a.n;
    d.mintb[i] = (int8_t)a.mintb;
  }
  static constexpr bool kSymmetric = true;  // r < max(H_i, H_j)
  static __device__ __forceinline__ void identity(Acc &a) {
    a.ax = a.ay = a.az = a.udt = a.hdt = 0.f;
    a.n = 0;
    a.mintb = kNumTimeBins + 1;
  }
  // hydro_reset_acceleration's sums and timestep_limiter_prepare_force's
  // bin (hydro.h:897-912)
  static __device__ __forceinline__ void fresh_acc(const I &, Acc &a) {
    identity(a);
  }
  template <int S>
  static __device__ __forceinline__ void combine(Acc &a) {
#pragma unroll
    for (int off = 64 / S; off < 64; off <<= 1) {
      a.ax += __shfl_xor(a.ax, off);
      a.ay += __shfl_xor(a.ay, off);
      a.az += __shfl_xor(a.az, off);
      a.udt += __shfl_xor(a.udt, off);
      a.hdt += __shfl_xor(a.hdt, off);
      a.n += __shfl_xor(a.n, off);
      a.mintb = min(a.mintb, __shfl_xor(a.mintb, off));
    }
  }
  // Symmetric criterion r < max(H_i, H_j) (DOPAIR2 :1915-2167, DOSELF2).
  template <class R>
  static __device__ __forceinline__ bool test(float r2, const I &s, const R *j) {
    return (r2 < s.hg2) || (r2 < j->hg2);
  }
  // runner_iact_nonsym_force (hydro_iact.h:507-632) + runner_iact_nonsym_
  // timebin (timestep_limiter_iact.h:59-70), a = 1, H = 0
  static __device__ __forceinline__ void interact(float r2, float dx, float dy,
                                                  float dz, const I &s,
                                                  const Rec &j, Acc &a) {
    const float r_inv = r2 > 0.f ? fast_rsq(r2) : 0.f;
    const float r = r2 * r_inv;
    float wi, wi_dx, wj, wj_dx;
    kernel_deval(r * s.h_inv, wi, wi_dx);
    kernel_deval(r * j.h_inv, wj, wj_dx);
    const float wi_dr = s.h_inv4 * wi_dx;
    const float hj2 = j.h_inv * j.h_inv;
    const float wj_dr = (hj2 * hj2) * wj_dx;
    const float dvdr =
        fmaf(s.vz - j.vz, dz, fmaf(s.vy - j.vy, dy, (s.vx - j.vx) * dx));
    const float mu = r_inv * fminf(dvdr, 0.f);
    const float v_sig = fmaf(-kViscBeta, mu, s.c + j.c);
    const float f_ij = fmaf(-s.f, j.inv_m, 1.f);
    const float f_ji = fmaf(-j.f, s.inv_m, 1.f);
    const float inv_rho_ij = fast_rcp(s.rho + j.rho);
    const float visc = -0.25f * (s.alpha + j.alpha) * v_sig * mu *
                       (s.bal + j.bal) * inv_rho_ij;
    const float visc_acc = 0.5f * visc * fmaf(wi_dr, f_ij, wj_dr * f_ji) * r_inv;
    const float Pi = s.pr2 * f_ij;
    const float Pj = j.pr2 * f_ji;
    const float acc = fmaf(fmaf(Pi, wi_dr, Pj * wj_dr), r_inv, visc_acc);
    const float mja = j.m * acc;
    a.ax = fmaf(-mja, dx, a.ax);
    a.ay = fmaf(-mja, dy, a.ay);
    a.az = fmaf(-mja, dz, a.az);
    const float dvdr_rinv = dvdr * r_inv;
    const float sph_du = Pi * dvdr_rinv * wi_dr;
    const float visc_du = 0.5f * visc_acc * dvdr;
    const float alpha_diff = (s.Pad + j.Pad) * fast_rcp(s.P + j.P);
    const float v_diff =
        alpha_diff * 0.5f *
        (fast_sqrt(2.f * fabsf(s.P - j.P) * inv_rho_ij) + fabsf(dvdr_rinv));
    const float diff_du = v_diff * (s.u - j.u) *
                          fmaf(f_ij * wi_dr, s.inv_rho, f_ji * wj_dr * j.inv_rho);
    a.udt = fmaf(sph_du + visc_du + diff_du, j.m, a.udt);
    a.hdt = fmaf(-j.m * dvdr_rinv * j.inv_rho, wi_dr, a.hdt);
    const int tbj = (int)j.tb;
    a.mintb = (tbj > 0 && tbj < a.mintb) ? tbj : a.mintb;
    a.n += 1;
  }
};

}  // namespace sphl
