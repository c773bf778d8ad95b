// sph_kernels.hip -- gfx950 kernels of the SPHENIX neighbour pass.
//
// Data layout (HBM, particles sorted by top-level cell):
//   px,py,pz double, vx..h float, tb int8: structure of arrays, read by the
//   per-particle kernels (ghost, extra ghost, end force);
//   DensRec / GradRec / ForceRec: one packed record per particle with the
//   fields a neighbour loop reads from particle j (sph_kernels.h), rebuilt
//   after the stage that changes them.
//
// The neighbour loops are in sph_gather.h (engine) and sph_loops.h (the
// three interaction bodies); this file holds the record builders, the
// per-particle kernels of the ghosts and the launchers.
#include <atomic>
#include "sph_device.h"
#include "sph_kernels.h"
#include "sph_loops.h"
#include "sph_gather.h"
#include "sph_sparse.h"

using namespace sph;
using namespace sphl;

namespace {

// ------------------------------------------------------ neighbour records --
// Thread = record slot j of a cell-aligned slot range (its particle is the
// field slot soa[j] with two slot orders, in the same cells).
__global__ void k_build_pos(DevParts d, GridDesc g, float pos_eps, int lo,
                            int hi) {
  const int j = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (j < hi) d.pos[j] = pos_rec(d, g, pos_eps, field_slot(d, j));
}

// Bounding box (non-inhibited particles) and largest search radius of each
// chunk in [c0, c1); an empty chunk gets an inverted box that no test passes.
// A wave takes 64 / C chunks, one lane per particle, and reduces over the C
// lanes of each chunk; the largest radius of the block goes to *hg2_all with
// one atomic per block.
constexpr int kBoundsBlock = 1024;
__global__ __launch_bounds__(kBoundsBlock) void k_chunk_bounds(
    DevParts d, const int *chunk_start, int c0, int c1, int C, double *bbox,
    float *hg2max, unsigned int *hg2_all, GridDesc gd, float pos_eps,
    float4 *fbox) {
  __shared__ float wmax[kBoundsBlock / 64];
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int k = c0 + wave * (64 / C) + lane / C;
  const int g = lane % C;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  float flo[3] = {3e38f, 3e38f, 3e38f}, fhi[3] = {-3e38f, -3e38f, -3e38f};
  float hm = 0.f, fh = -1.f;
  if (k < c1) {
    const int j = chunk_start[k] + g;
    const int i = j < chunk_start[k + 1] ? field_slot(d, j) : 0;
    if (j < chunk_start[k + 1] && d.tb[i] != kBinInhibited) {
      lo[0] = hi[0] = d.px[i];
      lo[1] = hi[1] = d.py[i];
      lo[2] = hi[2] = d.pz[i];
      hm = search_radius2(d.h[i]);
      if (fbox) {
        // the candidate's pre-test record, as k_build_pos forms it
        const PosRec p = pos_rec(d, gd, pos_eps, i);
        flo[0] = fhi[0] = p.x;
        flo[1] = fhi[1] = p.y;
        flo[2] = fhi[2] = p.z;
        fh = p.hg2m;
      }
    }
  }
  for (int off = 1; off < C; off <<= 1) {
    for (int a = 0; a < 3; a++) {
      lo[a] = fmin(lo[a], __shfl_xor(lo[a], off));
      hi[a] = fmax(hi[a], __shfl_xor(hi[a], off));
    }
    hm = fmaxf(hm, __shfl_xor(hm, off));
    if (fbox) {
      for (int a = 0; a < 3; a++) {
        flo[a] = fminf(flo[a], __shfl_xor(flo[a], off));
        fhi[a] = fmaxf(fhi[a], __shfl_xor(fhi[a], off));
      }
      fh = fmaxf(fh, __shfl_xor(fh, off));
    }
  }
  if (k < c1 && g == 0) {
    for (int a = 0; a < 3; a++) {
      bbox[6 * k + a] = lo[a];
      bbox[6 * k + 3 + a] = hi[a];
    }
    hg2max[k] = hm;
    if (fbox) {
      fbox[2 * k] = make_float4(flo[0], flo[1], flo[2], fh);
      fbox[2 * k + 1] = make_float4(fhi[0], fhi[1], fhi[2], 0.f);
    }
  }
  for (int off = 32; off >= C; off >>= 1) hm = fmaxf(hm, __shfl_xor(hm, off));
  if (lane == 0) wmax[threadIdx.x >> 6] = hm;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBoundsBlock / 64; k++) hm = fmaxf(hm, wmax[k]);
    // positive floats order like their bit patterns
    // a plain read first: only a block that raises the maximum pays an
    // atomic (one per block on one address serialised the whole launch)
    if (__float_as_uint(hm) >
        __hip_atomic_load(hg2_all, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(hg2_all, __float_as_uint(hm));
  }
}

// The radius part of k_chunk_bounds for chunks whose positions have not
// moved since their boxes were formed: the largest search radius (and, with
// fbox, the largest widened radius in fbox[2 k].w) of chunks [c0, c1), and
// with pos the widened radius of each member's pre-test record, as
// k_build_pos forms it.
__global__ __launch_bounds__(kBoundsBlock) void k_radius_bounds(
    DevParts d, const int *chunk_start, int c0, int c1, int C, float *hg2max,
    unsigned int *hg2_all, float pos_eps, float4 *fbox, PosRec *pos) {
  __shared__ float wmax[kBoundsBlock / 64];
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int k = c0 + wave * (64 / C) + lane / C;
  const int g = lane % C;
  float hm = 0.f, fh = -1.f;
  if (k < c1) {
    const int j = chunk_start[k] + g;
    if (j < chunk_start[k + 1]) {
      const int i = field_slot(d, j);
      const bool inh = d.tb[i] == kBinInhibited;
      if (!inh) {
        hm = search_radius2(d.h[i]);
        fh = pretest_radius2(hm, pos_eps);
      }
      if (pos) pos[j].hg2m = inh ? -1.f : fh;
    }
  }
  for (int off = 1; off < C; off <<= 1) {
    hm = fmaxf(hm, __shfl_xor(hm, off));
    fh = fmaxf(fh, __shfl_xor(fh, off));
  }
  if (k < c1 && g == 0) {
    hg2max[k] = hm;
    if (fbox) fbox[2 * k].w = fh;
  }
  for (int off = 32; off >= C; off >>= 1) hm = fmaxf(hm, __shfl_xor(hm, off));
  if (lane == 0) wmax[threadIdx.x >> 6] = hm;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k2 = 1; k2 < kBoundsBlock / 64; k2++) hm = fmaxf(hm, wmax[k2]);
    // a plain read first: only a block that raises the maximum pays an
    // atomic (one per block on one address serialised the whole launch)
    if (__float_as_uint(hm) >
        __hip_atomic_load(hg2_all, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(hg2_all, __float_as_uint(hm));
  }
}

// Records of particles [lo, hi); with max_active_bin >= 0 only of the
// active ones (the others keep theirs), stored one by one.  For all
// particles of the range, each wave stages its 64 records in LDS and writes
// them out as contiguous 16-byte words (a record per lane would scatter
// every store instruction over the wave's whole span).
constexpr int kRecBlock = 256;

// Store words of a record type: 16 bytes, or 8 for records of 8-byte
// multiples.
template <class R>
using RecWord = typename std::conditional<sizeof(R) % 16 == 0, uint4,
                                          uint2>::type;

template <class R>
__device__ __forceinline__ void store_wave_records(R *dst, int i0, int cnt,
                                                   const R &r,
                                                   RecWord<R> *lds) {
  using W = RecWord<R>;
  constexpr int Q = (int)(sizeof(R) / sizeof(W));
  static_assert(sizeof(R) % 8 == 0, "records are whole 8-byte words");
  const int lane = threadIdx.x & 63;
  const W *src = reinterpret_cast<const W *>(&r);
#pragma unroll
  for (int q = 0; q < Q; q++) lds[lane * Q + q] = src[q];
  __syncthreads();
  W *out = reinterpret_cast<W *>(dst + i0);
  const int words = cnt * Q;
  for (int k = lane; k < words; k += 64) out[k] = lds[k];
}

// LDS written by lanes of one wave and read by others of the same wave: no
// block barrier (the waves of a block take different branches here).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class R, class Make>
__device__ __forceinline__ void build_records(const DevParts &d, R *dst,
                                              int lo, int hi, int mab,
                                              const int *list, const int *cnt,
                                              Make make) {
  using W = RecWord<R>;
  constexpr int Q = (int)(sizeof(R) / sizeof(W));
  __shared__ W buf[kRecBlock * sizeof(R) / sizeof(W)];
  if (mab >= 0) {  // uniform; launched with run_blocks(hi - lo, kRecBlock)
    if (list || d.cand) {  // record slots not contiguous: one by one
      auto one = [&](int i, int8_t tb) {
        if (is_active(tb, mab)) dst[rec_slot(d, i)] = make(i);
      };
      if (list)
        for_runs_list(list, cnt, d.tb, one);
      else
        for_runs(lo, hi, d.tb, one);
      return;
    }
    // kRuns runs of the block's width; a run whose 64 particles of a wave
    // are all active is staged in the wave's LDS and stored as contiguous
    // words, otherwise the active ones store their records one by one
    const int lane = threadIdx.x & 63;
    W *mine = buf + (threadIdx.x >> 6) * 64 * Q;
    const int base = lo + blockIdx.x * kRecBlock * kRuns + threadIdx.x;
    int8_t t[kRuns];
#pragma unroll
    for (int u = 0; u < kRuns; u++) {
      const int i = base + u * kRecBlock;
      t[u] = i < hi ? d.tb[i] : (int8_t)kBinInhibited;
    }
    for (int u = 0; u < kRuns; u++) {
      const int i = base + u * kRecBlock;
      const bool act = i < hi && is_active(t[u], mab);
      if (__ballot(act) == ~0ull) {  // uniform over the wave
        const R r = make(i);
        const W *src = reinterpret_cast<const W *>(&r);
#pragma unroll
        for (int q = 0; q < Q; q++) mine[lane * Q + q] = src[q];
        wave_lds_sync();
        W *out = reinterpret_cast<W *>(dst + (i - lane));
        for (int k = lane; k < 64 * Q; k += 64) out[k] = mine[k];
        wave_lds_sync();  // the next run reuses the wave's LDS
      } else if (act) {
        dst[i] = make(i);
      }
    }
    return;
  }
  // thread = record slot; with two slot orders its particle is the field
  // slot soa[j] (the same cells' slot range), so the stores stay contiguous
  const int j = lo + blockIdx.x * kRecBlock + threadIdx.x;
  R r{};
  if (j < hi) r = make(field_slot(d, j));
  const int w = threadIdx.x >> 6;
  const int i0 = lo + blockIdx.x * kRecBlock + (w << 6);
  store_wave_records(dst, i0, max(0, min(64, hi - i0)), r, buf + w * 64 * Q);
}

__global__ __launch_bounds__(kRecBlock) void k_build_dens_rec(
    DevParts d, int lo, int hi, int mab, const int *list, const int *cnt) {
  build_records(d, d.rd, lo, hi, mab, list, cnt,
                [&](int i) { return dens_rec(d, i); });
}

__global__ __launch_bounds__(kRecBlock) void k_build_grad_rec(
    DevParts d, int lo, int hi, int mab, const int *list, const int *cnt) {
  build_records(d, d.rg, lo, hi, mab, list, cnt,
                [&](int i) { return grad_rec(d, i); });
}

__global__ __launch_bounds__(kRecBlock) void k_build_force_rec(
    DevParts d, int lo, int hi, int mab, const int *list, const int *cnt) {
  build_records(d, d.rf, lo, hi, mab, list, cnt,
                [&](int i) { return force_rec(d, i); });
}

// ------------------------------------------------------ per-particle work --
// hydro_init_part (hydro.h:564-606) for active particles in [lo, hi); also
// resets the ghost bracket.
__global__ void k_init_density(DevParts d, int lo, int hi, int max_active_bin,
                               float h_max, const int *list, const int *cnt,
                               int light) {
  auto body = [&](int i, int8_t tb) {
    const bool act = is_active(tb, max_active_bin);
    d.redo[i] = act ? 1 : 0;
    d.lfb[i] = 0;
    if (!act) return;
    d.left[i] = 0.f;
    d.right[i] = h_max;
    if (light) return;
    d.rho[i] = 0.f;
    d.rho_dh[i] = 0.f;
    d.wc[i] = 0.f;
    d.wc_dh[i] = 0.f;
    d.divv[i] = 0.f;
    d.rotx[i] = 0.f;
    d.roty[i] = 0.f;
    d.rotz[i] = 0.f;
    d.nd[i] = 0;
    d.lapu[i] = 0.f;
  };
  if (list)
    for_runs_list(list, cnt, d.tb, body);
  else
    for_runs(lo, hi, d.tb, body);
}

// hydro_prepare_gradient + hydro_reset_gradient (hydro.h:671-750) on a
// particle whose density sums are final.
__device__ void prepare_gradient(const DevParts &d, int i, float h, float rho,
                                 float rho_dh, float wc, float wc_dh,
                                 float divv, float rx, float ry, float rz,
                                 const GhostArgs &ga) {
  const float curl_v = sqrtf(rx * rx + ry * ry + rz * rz);
  const float abs_div_v = fabsf(divv);
  const float u = d.u[i];
  const float P = kGasGm1 * u * rho;
  const float c = sqrtf(kGasGamma * P / rho);
  const float bal = abs_div_v / (abs_div_v + curl_v + 0.0001f * c / h);
  const float common = h * kDimInv / wc;
  float fgh;
  if (h > 0.9999f * ga.h_max) {
    fgh = 0.f;
  } else {
    const float gW = common * wc_dh;
    fgh = (gW < -0.9999f) ? 0.f : common * rho_dh / (1.f + gW);
  }
  d.P[i] = P;
  d.c[i] = c;
  d.f[i] = fgh;
  d.bal[i] = bal;
  const float alpha = d.alpha[i];
  // the gradient record, built here from the values at hand
  if (ga.grad_rec) d.rg[rec_slot(d, i)] = grad_rec_of(d, i, rho, u, c, alpha);
  if (ga.fresh_grad) return;  // the gradient loop starts from these itself
  d.vsig[i] = 2.f * c;
  d.avmax[i] = alpha;
  d.lapu[i] = 0.f;
  d.ng[i] = 0;
}

// One ghost round (runner_do_ghost, runner_ghost.c:1113-1628) for the
// particles flagged in redo[] within [lo, hi): hydro_end_density, one
// Newton-Raphson / bisection step on h, and either convergence (prepare the
// gradient loop) or a re-initialised particle flagged for another density
// round.  `relist` tells whether the particle's neighbour list no longer
// covers its new search radius (for another round, or for the gradient
// loop once converged).
__device__ bool ghost_one(const DevParts &d, const GhostArgs &ga, int i,
                          bool &relist) {
  const float h_old = d.h[i];
  const float h_old_dim = h_old * h_old * h_old;
  const float h_old_dim_m1 = h_old * h_old;
  float rho = d.rho[i], rho_dh = d.rho_dh[i], wc = d.wc[i], wc_dh = d.wc_dh[i];
  float divv = d.divv[i], rx = d.rotx[i], ry = d.roty[i], rz = d.rotz[i];
  float left = d.left[i], right = d.right[i];
  float h_new;
  bool no_ngb = false;
  bool done = false;
  if ((double)wc < 1.e-5 * (double)kRoot) {
    no_ngb = true;
    h_new = 2.f * h_old;
  } else {
    // hydro_end_density (hydro.h:613-662), a = 1, H = 0
    const float h_inv = 1.0f / h_old;
    const float hid = h_inv * h_inv * h_inv;
    const float hid1 = hid * h_inv;
    const float mi = d.m[i];
    rho = (rho + mi * kRoot) * hid;
    rho_dh = (rho_dh - kDim * mi * kRoot) * hid1;
    wc = (wc + kRoot) * hid;
    wc_dh = (wc_dh - kDim * kRoot) * hid1;
    const float rho_inv = 1.f / rho;
    rx *= hid1 * rho_inv;
    ry *= hid1 * rho_inv;
    rz *= hid1 * rho_inv;
    divv *= hid1 * rho_inv;
    const float n_sum = wc * h_old_dim;
    const float f = n_sum - ga.eta_dim;
    const float f_prime = wc_dh * h_old_dim + kDim * wc * h_old_dim_m1;
    if (n_sum < ga.eta_dim)
      left = fmaxf(left, h_old);
    else if (n_sum > ga.eta_dim)
      right = fminf(right, h_old);
    if (((h_old >= ga.h_max) && (f < 0.f)) ||
        ((h_old <= ga.h_min) && (f > 0.f))) {
      done = true;
      h_new = h_old;
    } else {
      h_new = h_old - f / (f_prime + 1.17549435e-38f);
      h_new = fminf(h_new, 2.f * h_old);
      h_new = fmaxf(h_new, 0.5f * h_old);
      h_new = fmaxf(h_new, left);
      h_new = fminf(h_new, right);
    }
  }
  float h_fin = h_old;
  if (!done && fabsf(h_new - h_old) > ga.eps * h_old) {
    float hn;
    if ((h_new == left && h_old == right) || (h_old == left && h_new == right))
      hn = cbrtf(0.5f * (left * left * left + right * right * right));
    else
      hn = h_new;
    if (hn < ga.h_max && hn > ga.h_min) {
      // Not converged: re-initialise (hydro_init_part) and ask for a rerun.
      d.h[i] = hn;
      d.left[i] = left;
      d.right[i] = right;
      d.rho[i] = 0.f;
      d.rho_dh[i] = 0.f;
      d.wc[i] = 0.f;
      d.wc_dh[i] = 0.f;
      d.divv[i] = 0.f;
      d.rotx[i] = 0.f;
      d.roty[i] = 0.f;
      d.rotz[i] = 0.f;
      d.nd[i] = 0;
      relist = !ga.lists || !(search_radius2(hn) <= d.r2b[i]);
      return true;
    }
    if (hn <= ga.h_min) {
      h_fin = ga.h_min;
    } else {
      h_fin = ga.h_max;
      if (no_ngb) {
        // hydro_part_has_no_neighbours (hydro.h:791-825)
        const float hinv = 1.0f / h_fin;
        const float hid = hinv * hinv * hinv;
        rho = d.m[i] * kRoot * hid;
        wc = kRoot * hid;
        rho_dh = 0.f;
        wc_dh = 0.f;
        rx = ry = rz = 0.f;
        divv = 0.f;
      }
    }
  }
  // Converged.
  d.h[i] = h_fin;
  d.redo[i] = 0;
  relist = !ga.lists || !(search_radius2(h_fin) <= d.r2b[i]);
  if (ga.lists) d.lfb[i] = relist ? 1 : 0;
  d.rho[i] = rho;
  d.rho_dh[i] = rho_dh;
  d.wc[i] = wc;
  d.wc_dh[i] = wc_dh;
  d.divv[i] = divv;
  d.rotx[i] = rx;
  d.roty[i] = ry;
  d.rotz[i] = rz;
  prepare_gradient(d, i, h_fin, rho, rho_dh, wc, wc_dh, divv, rx, ry, rz, ga);
  return false;
}

// Counts are reduced over the block first: one atomic per block and
// counter keeps the same-address traffic small.  A block covers kRuns runs
// of kGhostBlock particles (their redo flags loaded together); the
// particles that go again are appended to d.sel in block, wave, run, lane
// order.
constexpr int kGhostBlock = 256;
__global__ __launch_bounds__(kGhostBlock) void k_ghost(
    DevParts d, GhostArgs ga, int lo, int hi, int off, int nt,
    unsigned long long *counters, const int *list, const int *cnt) {
  constexpr int W = kGhostBlock / 64;
  __shared__ int wcnt[3][W];
  __shared__ unsigned long long wwork[W];
  __shared__ int base;
  // run u of this thread: particle slot i0 + u * kGhostBlock, or with a
  // compact list the particle of list entry k0 + u * kGhostBlock
  const int i0 = lo + blockIdx.x * kGhostBlock * kRuns + threadIdx.x;
  const int k0 = blockIdx.x * kGhostBlock * kRuns + threadIdx.x;
  const int nl = list ? *cnt : 0;
  int idx[kRuns], redo[kRuns];
#pragma unroll
  for (int u = 0; u < kRuns; u++) {
    if (list) {
      const int k = k0 + u * kGhostBlock;
      idx[u] = k < nl ? list[k] : -1;
    } else {
      // share off of nt of the slots (runner_do_ghost's offset / ntasks)
      const int i = i0 + u * kGhostBlock;
      idx[u] = (i < hi && (nt == 1 || (i - lo) % nt == off)) ? i : -1;
    }
  }
#pragma unroll
  for (int u = 0; u < kRuns; u++) redo[u] = idx[u] >= 0 ? d.redo[idx[u]] : 0;
  unsigned int again = 0u, relist = 0u;  // bit u: run u
  // density interactions evaluated for these particles in the last round
  unsigned long long work = 0ull;
#pragma unroll
  for (int u = 0; u < kRuns; u++) {
    if (redo[u]) {
      const int i = idx[u];
      work += (unsigned long long)d.nd[i];
      bool rl = false;
      if (ghost_one(d, ga, i, rl)) again |= 1u << u;
      if (rl) relist |= 1u << u;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) work += __shfl_xor(work, off);
  int na = 0, nb = 0, nc = 0;
#pragma unroll
  for (int u = 0; u < kRuns; u++) {
    const bool ag = (again >> u) & 1u, rl = (relist >> u) & 1u;
    na += __popcll(__ballot(ag));
    nb += __popcll(__ballot(ag && rl));
    nc += ga.lists ? __popcll(__ballot(!ag && rl)) : 0;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    wcnt[0][w] = na;
    wcnt[1][w] = nb;
    wcnt[2][w] = nc;
    wwork[w] = work;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long t = 0;
    for (int k = 0; k < W; k++) t += (unsigned long long)wcnt[threadIdx.x][k];
    unsigned long long b0 = 0;
    if (t) b0 = atomicAdd(&counters[threadIdx.x], t);
    if (threadIdx.x == 0) base = (int)b0;
  } else if (threadIdx.x == 3) {
    unsigned long long t = 0;
    for (int k = 0; k < W; k++) t += wwork[k];
    if (t) atomicAdd(&counters[3 + blockIdx.x % kGhostWorkSlots], t);
  }
  __syncthreads();
  if (__any(again != 0u)) {
    // append to the compact list of particles for the next round
    int off = base;
    for (int k = 0; k < w; k++) off += wcnt[0][k];
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int u = 0; u < kRuns; u++) {
      const bool ag = (again >> u) & 1u;
      const unsigned long long ba = __ballot(ag);
      if (ag) d.sel[off + __popcll(ba & below)] = idx[u];
      off += __popcll(ba);
    }
  }
}

__global__ void k_fill_f32(float *p, float v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// runner_do_extra_ghost (runner_ghost.c:1016-1111): hydro_end_gradient,
// hydro_prepare_force (hydro.h:840-954, a = 1), hydro_reset_acceleration and
// timestep_limiter_prepare_force for active particles in [lo, hi).
__global__ void k_extra_ghost(DevParts d, GhostArgs ga, int lo, int hi,
                              const int *list, const int *cnt) {
  auto body = [&](int i, int8_t tb) {
    if (!is_active(tb, ga.max_active_bin)) return;
    const float h = d.h[i];
    const float h_inv = 1.0f / h;
    const float hid1 = h_inv * h_inv * h_inv * h_inv;
    const float lapu = d.lapu[i] * (2.f * hid1);
    d.lapu[i] = lapu;
    const float dt_alpha =
        tb <= 0 ? 0.f : (float)((double)(1LL << (tb + 1)) * ga.time_base);
    const float ksp = h * kGamma;
    const float ksp_inv = 1.f / ksp;
    const float vsig = d.vsig[i];
    const float rho = d.rho[i];
    const float P = kGasGm1 * d.u[i] * rho;
    const float c = sqrtf(kGasGamma * P / rho);
    const float sct_inv = c * ksp_inv;
    const float divv = d.divv[i];
    const float div_v_dt =
        dt_alpha == 0.f ? 0.f : (divv - d.divv_prev[i]) / dt_alpha;
    const float S = divv < 0.f ? ksp * ksp * fmaxf(0.f, -div_v_dt) : 0.f;
    const float alpha_loc = ga.alpha_max * S / (c * c + S);
    float alpha = d.alpha[i];
    if (alpha_loc > alpha) {
      alpha = alpha_loc;
    } else {
      const float tr = dt_alpha * sct_inv * ga.visc_length;
      alpha += alpha_loc * tr;
      alpha /= (1.f + tr);
    }
    alpha = fmaxf(alpha, ga.alpha_min);
    d.alpha[i] = alpha;
    d.divv_prev[i] = divv;
    d.divv_dt[i] = div_v_dt;
    const float dtpi = vsig * ksp_inv;
    const float sqrt_u_inv = 1.f / sqrtf(d.u[i]);
    float adt = ga.diff_beta * ksp * lapu * sqrt_u_inv;
    const float ad0 = d.adiff[i];
    adt -= (ad0 - ga.diff_alpha_min) * dtpi;
    float na = ad0 + adt * dt_alpha;
    na = fmaxf(na, ga.diff_alpha_min);
    const float lim = ga.diff_alpha_max * (1.f - d.avmax[i] / ga.alpha_max);
    d.adiff[i] = fminf(na, lim);
    // the force record, built here from the values at hand
    if (ga.force_rec)
      d.rf[rec_slot(d, i)] = force_rec_of(d, i, tb, h, rho, P, c, alpha,
                                          fminf(na, lim), d.u[i]);
    if (ga.fresh_force) return;
    d.ax[i] = 0.f;
    d.ay[i] = 0.f;
    d.az[i] = 0.f;
    d.udt[i] = 0.f;
    d.hdt[i] = 0.f;
    d.nf[i] = 0;
    d.mintb[i] = (int8_t)(kNumTimeBins + 1);
  };
  if (list)
    for_runs_list(list, cnt, d.tb, body);
  else
    for_runs(lo, hi, d.tb, body);
}

// runner_do_end_hydro_force (runner_others.c:815-870) -> hydro_end_force.
__global__ void k_end_force(DevParts d, int lo, int hi, int max_active_bin,
                            const int *list, const int *cnt) {
  auto body = [&](int i, int8_t tb) {
    if (is_active(tb, max_active_bin)) d.hdt[i] *= d.h[i] * kDimInv;
  };
  if (list)
    for_runs_list(list, cnt, d.tb, body);
  else
    for_runs(lo, hi, d.tb, body);
}

__global__ void k_zero_int(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// mask[c] = 1 for the n cells of `cells`, 0 for the others (two launches:
// clear, then set).
__global__ void k_zero_u8(unsigned char *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_set_mask(unsigned char *mask, const int *cells, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) mask[cells[k]] = 1;
}

// act[c] = 1 for the cells holding an active particle (cleared before).
__global__ void k_cell_active(DevParts d, int n, int max_active_bin,
                              unsigned char *act) {
  for_runs(0, n, d.tb, [&](int i, int8_t tb) {
    if (is_active(tb, max_active_bin)) act[d.cell[i]] = 1;
  });
}

// inv[perm[s]] = s: device slot of every caller index.
__global__ void k_invert_perm(const int *perm, int *inv, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) inv[perm[s]] = s;
}

// *moved = 1 when perm is not the identity (flag zeroed first).
__global__ void k_perm_moved(const int *perm, int n, unsigned int *moved) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n && perm[s] != s) *moved = 1u;
}

// Chunk table of capacity C from the cells' slot ranges: chunk ch of cell c
// (cell_chunk[c] <= ch < cell_chunk[c + 1]) starts at slot cell_start[c] +
// C * (ch - cell_chunk[c]); start[nchunks] = n closes the last one.  The
// cell of a chunk is found by bisection over cell_chunk.
__global__ void k_chunk_fill(const int *cell_start, const int *cell_chunk,
                             int ncells, int C, int nchunks, int n,
                             int *start, int *chunk_cell) {
  const int ch = blockIdx.x * blockDim.x + threadIdx.x;
  if (ch > nchunks) return;
  if (ch == nchunks) {
    start[ch] = n;
    return;
  }
  int lo = 0, hi = ncells;  // cell_chunk[lo] <= ch < cell_chunk[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cell_chunk[mid] <= ch)
      lo = mid;
    else
      hi = mid;
  }
  start[ch] = cell_start[lo] + C * (ch - cell_chunk[lo]);
  chunk_cell[ch] = lo;
}

// flag[slot of caller index first + ind[k]] = 1 for k < count; a particle
// whose slot lies outside [lo, hi) (not in the task's cell) is counted in
// *err instead.  The caller checked first + ind[k] against the particle
// count.
__global__ void k_mark_subset(int *flag, const int *inv, long long first,
                              const int *ind, int count, int lo, int hi,
                              unsigned int *err) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int s = inv[first + ind[k]];
  if (s >= lo && s < hi)
    flag[s] = 1;
  else
    atomicAdd(err, 1u);
}

// dst[i] = src[i] for the particles i of [lo, hi) with (i - lo) % nt == off,
// 0 for the others: the share of a cell's ghost task `off` of `nt`.
__global__ void k_share_flags(int *dst, const int *src, int lo, int hi,
                              int off, int nt) {
  const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < hi) dst[i] = ((i - lo) % nt == off) ? src[i] : 0;
}

// Sums the per-particle interaction counters of active particles:
// out[0] density, out[1] gradient, out[2] force.  One atomic per block.
__global__ void k_count(DevParts d, int n, int max_active_bin,
                        unsigned long long *out) {
  __shared__ unsigned long long part[3][4];
  unsigned long long a = 0, b = 0, c = 0;
  // kRuns time bins in flight per thread and sweep
  const int stride = gridDim.x * blockDim.x * kRuns;
  for (int i0 = blockIdx.x * blockDim.x * kRuns + threadIdx.x; i0 < n;
       i0 += stride) {
    int8_t t[kRuns];
#pragma unroll
    for (int u = 0; u < kRuns; u++) {
      const int i = i0 + u * (int)blockDim.x;
      t[u] = i < n ? d.tb[i] : kBinInhibited;
    }
#pragma unroll
    for (int u = 0; u < kRuns; u++) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < n && is_active(t[u], max_active_bin)) {
        a += (unsigned)d.nd[i];
        b += (unsigned)d.ng[i];
        c += (unsigned)d.nf[i];
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off);
    b += __shfl_down(b, off);
    c += __shfl_down(c, off);
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][wv] = a;
    part[1][wv] = b;
    part[2][wv] = c;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long s = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += part[threadIdx.x][k];
    atomicAdd(&out[threadIdx.x], s);
  }
}

// Block (x, k) copies 1024 floats of array k as float4 words (the buffers
// are 256-byte aligned; the last partial word element by element).
__global__ __launch_bounds__(256) void k_copy_set(sphk::CopySet cs, int n) {
  const int k = blockIdx.y;
  const int w = blockIdx.x * 256 + threadIdx.x;  // float4 word
  const int i = 4 * w;
  if (i + 3 < n) {
    reinterpret_cast<float4 *>(cs.dst[k])[w] =
        reinterpret_cast<const float4 *>(cs.src[k])[w];
  } else {
    for (int j = i; j < n; j++) cs.dst[k][j] = cs.src[k][j];
  }
}

// Active particles of [0, n) at max_active_bin: one atomic per block.
__global__ void k_count_active(const int8_t *__restrict__ tb, int n,
                               int max_active_bin, unsigned int *out) {
  __shared__ unsigned int part[4];
  unsigned int c = 0;
  const int stride = gridDim.x * blockDim.x * kRuns;
  for (int i0 = blockIdx.x * blockDim.x * kRuns + threadIdx.x; i0 < n;
       i0 += stride) {
    int8_t t[kRuns];
#pragma unroll
    for (int u = 0; u < kRuns; u++) {
      const int i = i0 + u * (int)blockDim.x;
      t[u] = i < n ? tb[i] : kBinInhibited;
    }
#pragma unroll
    for (int u = 0; u < kRuns; u++) c += is_active(t[u], max_active_bin) ? 1u : 0u;
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int t = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += part[k];
    if (t) atomicAdd(out, t);
  }
}

// ------------------------------------------------------------ tile builder --
// Device-wide exclusive scans (block sums, one-block scan of the sums,
// block-local scans) and the compaction / tiling of the particles a loop
// processes.  Everything stays on the device.
constexpr int kScanBlock = 1024;

// Exclusive scan of v over the block; *total receives the block sum.
__device__ __forceinline__ int block_excl_scan(int v, int *wsum, int *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (w == 0) {
    const int x = lane < kScanBlock / 64 ? wsum[lane] : 0;
    int xi = x;
#pragma unroll
    for (int off = 1; off < kScanBlock / 64; off <<= 1) {
      const int t = __shfl_up(xi, off);
      if (lane >= off) xi += t;
    }
    if (lane < kScanBlock / 64) wsum[lane] = xi - x;
    if (lane == kScanBlock / 64 - 1) wsum[kScanBlock / 64] = xi;
  }
  __syncthreads();
  *total = wsum[kScanBlock / 64];
  return wsum[w] + incl - v;
}

__device__ __forceinline__ bool tile_pred(const DevParts &d, const LoopArgs &la,
                                          int mode, const int *flag, int i,
                                          int8_t tb) {
  if (!is_active(tb, la.max_active_bin)) return false;
  if (la.cmask && !la.cmask[d.cell[i]]) return false;
  if (mode >= 1 && !flag[i]) return false;
  // mode 2: only particles whose neighbour list does not cover their radius
  if (mode == 2 && search_radius2(d.h[i]) <= d.r2b[i]) return false;
  return true;
}

// f[s] = 1 for the particles the tiles take, in slot order s.  With two
// slot orders the slots are record slots (f[j] for the particle in field
// slot soa[j]), so that the particles of a cell that several active time
// bins hold still form tiles in the cell's Morton order; with one active
// bin both orders list them alike.
__global__ void k_sel_flag(DevParts d, LoopArgs la, int mode, const int *flag,
                           int n, int *f) {
  if (d.soa) {
    const int base = blockIdx.x * blockDim.x * kRuns + threadIdx.x;
    int ii[kRuns];
#pragma unroll
    for (int u = 0; u < kRuns; u++) {
      const int j = base + u * (int)blockDim.x;
      ii[u] = j < n ? d.soa[j] : -1;
    }
    int8_t tb[kRuns];
#pragma unroll
    for (int u = 0; u < kRuns; u++) tb[u] = ii[u] >= 0 ? d.tb[ii[u]] : 0;
#pragma unroll
    for (int u = 0; u < kRuns; u++)
      if (ii[u] >= 0)
        f[base + u * (int)blockDim.x] =
            tile_pred(d, la, mode, flag, ii[u], tb[u]) ? 1 : 0;
    return;
  }
  for_runs(0, n, d.tb, [&](int i, int8_t tb) {
    f[i] = tile_pred(d, la, mode, flag, i, tb) ? 1 : 0;
  });
}

__global__ __launch_bounds__(kScanBlock) void k_scan_up(const int *in, int n,
                                                        int *blk) {
  __shared__ int wsum[kScanBlock / 64 + 1];
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  int total;
  (void)block_excl_scan(i < n ? in[i] : 0, wsum, &total);
  if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// One block: exclusive scan of blk[0..nb) in place, the total to *total_out.
__global__ __launch_bounds__(kScanBlock) void k_scan_blocks(int *blk, int nb,
                                                            int *total_out) {
  __shared__ int wsum[kScanBlock / 64 + 1];
  const int per = (nb + kScanBlock - 1) / kScanBlock;
  const int a = threadIdx.x * per, b = min(nb, a + per);
  int sum = 0;
  for (int k = a; k < b; k++) sum += blk[k];
  int total;
  int run = block_excl_scan(sum, wsum, &total);
  for (int k = a; k < b; k++) {
    const int v = blk[k];
    blk[k] = run;
    run += v;
  }
  if (threadIdx.x == 0) *total_out = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(const int *in, int n,
                                                          const int *blk,
                                                          int *out) {
  __shared__ int wsum[kScanBlock / 64 + 1];
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  int total;
  const int ex = block_excl_scan(i < n ? in[i] : 0, wsum, &total);
  if (i < n) out[i] = blk[blockIdx.x] + ex;
}

// alist[S[s]] = the field slot of slot s (soa[s] with two slot orders).
__global__ void k_sel_scatter(const int *f, const int *S, int n,
                              const int *soa, int *alist) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n && f[s]) alist[S[s]] = soa ? soa[s] : s;
}

__global__ void k_tile_count(const int *cell_start, const int *S, int ncells,
                             int C, int *cnt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int k = S[cell_start[c + 1]] - S[cell_start[c]];
  cnt[c] = (k + C - 1) / C;
}

__global__ void k_tile_fill(const int *cell_start, const int *S, const int *T,
                            int ncells, int C, int n, int *tile_start,
                            int *tile_cell, int *ntiles) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) {
    const int base = S[cell_start[c]];
    const int k = S[cell_start[c + 1]] - base;
    const int t0 = T[c];
    for (int t = 0; t * C < k; t++) {
      tile_start[t0 + t] = base + t * C;
      tile_cell[t0 + t] = c;
    }
  }
  if (c == 0) {
    const int nt = T[ncells];
    tile_start[nt] = S[n];
    *ntiles = nt;
  }
}

// Tiles with a particle of filt[i] != 0, appended to ids with one atomic
// per wave (grid-stride over the device-side tile count).
__global__ void k_select_tiles(const int *tile_start, const int *alist,
                               const int *ntiles, const int *filt, int *ids,
                               int *nids) {
  const int nt = *ntiles;
  const int lane = threadIdx.x & 63;
  for (int b = blockIdx.x * blockDim.x; b < nt; b += gridDim.x * blockDim.x) {
    const int t = b + threadIdx.x;
    bool any = false;
    if (t < nt) {
      const int e = tile_start[t + 1];
      for (int k = tile_start[t]; k < e && !any; k++) any = filt[alist[k]] != 0;
    }
    const unsigned long long m = __ballot(any);
    int off = 0;
    if (lane == 0 && m) off = atomicAdd(nids, __popcll(m));
    off = __shfl(off, 0);
    if (any)
      ids[off + __popcll(m & ((1ull << lane) - 1ull))] = t;
  }
}

}  // namespace

// --------------------------------------------------------------- launchers --
namespace sphk {

static inline int nblk(int n, int b) { return (n + b - 1) / b; }
// Blocks (one wave each) of the grid that takes the tiles beyond a captured
// grid's reach.
constexpr int kRestBlocks = 2048;

template <class L, bool SUB, int C>
void stencil_c(const DevParts &d, const LoopArgs &la, const ChunkArgs &ca,
               const typename L::Rec *rec, hipStream_t st) {
  const NbrList none{};
  hipLaunchKernelGGL((k_loop_stencil<L, C, SUB, false>),
                     dim3(xcd_blocks(ca.nchunks)), dim3(64 * kWaves), 0, st, d,
                     la, ca, rec, (const PosRec *)d.pos, none);
}

hipError_t launch_build_lists(bool subset, const DevParts &d,
                              const LoopArgs &la, const ChunkArgs &ca,
                              const NbrList &nl, hipStream_t st) {
  if (ca.nchunks <= 0) return hipSuccess;
  if (ca.C != kListChunk) return hipErrorInvalidValue;
  const dim3 grid(xcd_blocks(ca.nchunks)), blk(64 * kWaves);
  if (subset)
    hipLaunchKernelGGL((k_loop_stencil<DensityLoop, kListChunk, true, true>),
                       grid, blk, 0, st, d, la, ca, d.rd,
                       (const PosRec *)d.pos, nl);
  else
    hipLaunchKernelGGL((k_loop_stencil<DensityLoop, kListChunk, false, true>),
                       grid, blk, 0, st, d, la, ca, d.rd,
                       (const PosRec *)d.pos, nl);
  return hipGetLastError();
}

template <class L>
void list_loop(bool compact, const DevParts &d, const LoopArgs &la,
               const ChunkArgs &ca, const NbrList &nl,
               const typename L::Rec *rec, int nsel, hipStream_t st) {
  if (compact) {
    hipLaunchKernelGGL((k_loop_list<L, true>),
                       dim3(nblk(nsel, kListChunk * kWaves)), dim3(64 * kWaves),
                       0, st, d, la, ca, nl, rec, nsel);
  } else {
    hipLaunchKernelGGL((k_loop_list<L, false>), dim3(xcd_blocks(ca.nchunks)),
                       dim3(64 * kWaves), 0, st, d, la, ca, nl, rec, 0);
  }
}

hipError_t launch_list_loop(LoopKind kind, bool compact, const DevParts &d,
                            const LoopArgs &la, const ChunkArgs &ca,
                            const NbrList &nl, int nsel, hipStream_t st) {
  if (compact ? nsel <= 0 : ca.nchunks <= 0) return hipSuccess;
  if (!compact && ca.C != kListChunk) return hipErrorInvalidValue;
  switch (kind) {
    case LoopKind::Density:
      list_loop<DensityLoop>(compact, d, la, ca, nl, d.rd, nsel, st);
      break;
    case LoopKind::Gradient:
      list_loop<GradientLoop>(compact, d, la, ca, nl, d.rg, nsel, st);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_build_tiles(const DevParts &d, const LoopArgs &la, int n,
                              int ncells, int C, int mode, const int *flag,
                              const TileScratch &ts, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_int, dim3(8), dim3(64), 0, st, ts.heads, 8 * 64);
  if (n <= 0) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, ts.ntiles, 1);
    return hipGetLastError();
  }
  const int nb = nblk(n, kScanBlock), ncb = nblk(ncells, kScanBlock);
  hipLaunchKernelGGL(k_sel_flag, dim3(run_blocks(n, 256)), dim3(256), 0, st, d, la,
                     mode, flag, n, ts.flag);
  hipLaunchKernelGGL(k_scan_up, dim3(nb), dim3(kScanBlock), 0, st,
                     (const int *)ts.flag, n, ts.blk);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanBlock), 0, st, ts.blk,
                     nb, ts.scan + n);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanBlock), 0, st,
                     (const int *)ts.flag, n, (const int *)ts.blk, ts.scan);
  hipLaunchKernelGGL(k_sel_scatter, dim3(nblk(n, 256)), dim3(256), 0, st,
                     (const int *)ts.flag, (const int *)ts.scan, n,
                     (const int *)d.soa, ts.alist);
  hipLaunchKernelGGL(k_tile_count, dim3(nblk(ncells, 256)), dim3(256), 0, st,
                     la.cell_start, (const int *)ts.scan, ncells, C, ts.cnt);
  hipLaunchKernelGGL(k_scan_up, dim3(ncb), dim3(kScanBlock), 0, st,
                     (const int *)ts.cnt, ncells, ts.cblk);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanBlock), 0, st, ts.cblk,
                     ncb, ts.cscan + ncells);
  hipLaunchKernelGGL(k_scan_down, dim3(ncb), dim3(kScanBlock), 0, st,
                     (const int *)ts.cnt, ncells, (const int *)ts.cblk,
                     ts.cscan);
  hipLaunchKernelGGL(k_tile_fill, dim3(nblk(ncells, 256)), dim3(256), 0, st,
                     la.cell_start, (const int *)ts.scan, (const int *)ts.cscan,
                     ncells, C, n, ts.tile_start, ts.tile_cell, ts.ntiles);
  return hipGetLastError();
}

// Resident grid of a persistent kernel: blocks per CU at full occupancy
// times the CU count (queried once per kernel instance).
template <class L, int CI, int CJ, bool B>
static int resident_blocks() {
  static std::atomic<int> cached{0};
  if (cached.load() == 0) {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, k_loop_queue<L, CI, CJ, B>, 64 * kWaves, 0) != hipSuccess ||
        per <= 0)
      return 0;
    cached.store(ncu * per * kWaves);
  }
  return cached.load();
}

template <class L, int CI, int CJ, bool B>
hipError_t queue_c(const DevParts &d, const LoopArgs &la, const ChunkArgs &ca,
                   const TileArgs &ta, const NbrList &nl,
                   const typename L::Rec *rec, bool persistent, int waves,
                   int max_tiles, bool rest, hipStream_t st) {
  if (!persistent) {
    const int nb = max_tiles > 0 ? xcd_blocks(max_tiles) : 0;
    if (nb > 0)
      hipLaunchKernelGGL((k_loop_tiles<L, CI, CJ, B>), dim3(nb),
                         dim3(64 * kWaves), 0, st, d, la, ca, rec,
                         (const PosRec *)d.pos, nl, ta);
    if (rest)
      hipLaunchKernelGGL((k_loop_tiles_rest<L, CI, CJ, B>), dim3(kRestBlocks),
                         dim3(64), 0, st, d, la, ca, rec,
                         (const PosRec *)d.pos, nl, ta, nb);
    return hipGetLastError();
  }
  auto kern = k_loop_queue<L, CI, CJ, B>;
  if (waves <= 0) waves = resident_blocks<L, CI, CJ, B>();
  if (waves <= 0) return hipErrorInvalidValue;
  // `waves` counts waves; blocks hold kWaves of them
  hipLaunchKernelGGL(kern, dim3((waves + kWaves - 1) / kWaves),
                     dim3(64 * kWaves), 0, st, d, la, ca, rec,
                     (const PosRec *)d.pos, nl, ta);
  return hipGetLastError();
}

// Tile (CI) x chunk (CJ) shapes with kernels: CJ = CI, CJ = CI / 2 for
// CI = 16, 32 (finer candidate chunks under the same tile), and the narrow
// tiles (4, 16), (4, 32), (8, 16) for sparse activity (a few active particles
// per cell: more candidate streams per particle, fewer idle lanes).
template <class L>
hipError_t queue_any(const DevParts &d, const LoopArgs &la, const ChunkArgs &ca,
                     const TileArgs &ta, const NbrList &nl,
                     const typename L::Rec *rec, int CI, int CJ, bool p,
                     int waves, int mt, bool rest, hipStream_t st) {
  switch (CI * 100 + CJ) {
    case 808: return queue_c<L, 8, 8, false>(d, la, ca, ta, nl, rec, p, waves, mt, rest, st);
    case 1616: return queue_c<L, 16, 16, false>(d, la, ca, ta, nl, rec, p, waves, mt, rest, st);
    case 1608: return queue_c<L, 16, 8, false>(d, la, ca, ta, nl, rec, p, waves, mt, rest, st);
    case 3232: return queue_c<L, 32, 32, false>(d, la, ca, ta, nl, rec, p, waves, mt, rest, st);
    case 3216: return queue_c<L, 32, 16, false>(d, la, ca, ta, nl, rec, p, waves, mt, rest, st);
    case 6464: return queue_c<L, 64, 64, false>(d, la, ca, ta, nl, rec, p, waves, mt, rest, st);
    case 416: return queue_c<L, 4, 16, false>(d, la, ca, ta, nl, rec, p, waves, mt, rest, st);
    case 432: return queue_c<L, 4, 32, false>(d, la, ca, ta, nl, rec, p, waves, mt, rest, st);
    case 816: return queue_c<L, 8, 16, false>(d, la, ca, ta, nl, rec, p, waves, mt, rest, st);
    case 1632: return queue_c<L, 16, 32, false>(d, la, ca, ta, nl, rec, p, waves, mt, rest, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_queue(LoopKind kind, bool build_lists, const DevParts &d,
                        const LoopArgs &la, const ChunkArgs &ca,
                        const TileArgs &ta, const NbrList &nl, int CI, int CJ,
                        bool persistent, int waves, int max_tiles, bool rest,
                        hipStream_t st) {
  if (ca.C != CJ) return hipErrorInvalidValue;
  switch (kind) {
    case LoopKind::Density:
      if (build_lists) {
        if (CI != kListChunk || CJ != kListChunk) return hipErrorInvalidValue;
        return queue_c<DensityLoop, kListChunk, kListChunk, true>(
            d, la, ca, ta, nl, d.rd, persistent, waves, max_tiles, rest, st);
      }
      return queue_any<DensityLoop>(d, la, ca, ta, nl, d.rd, CI, CJ,
                                    persistent, waves, max_tiles, rest,
                                    st);
    case LoopKind::Gradient:
      return queue_any<GradientLoop>(d, la, ca, ta, nl, d.rg, CI, CJ,
                                     persistent, waves, max_tiles, rest,
                                    st);
    case LoopKind::Force:
      return queue_any<ForceLoop>(d, la, ca, ta, nl, d.rf, CI, CJ,
                                  persistent, waves, max_tiles, rest,
                                    st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_select_tiles(const TileScratch &ts, const int *filt,
                               int *ids, int *nids, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, nids, 1);
  hipLaunchKernelGGL(k_select_tiles, dim3(1024), dim3(256), 0, st,
                     (const int *)ts.tile_start, (const int *)ts.alist,
                     (const int *)ts.ntiles, filt, ids, nids);
  return hipGetLastError();
}

template <class L>
static int sparse_resident_blocks() {
  static std::atomic<int> cached{0};
  if (cached.load() == 0) {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_sparse_queue<L>,
                                                     64, 0) != hipSuccess ||
        per <= 0)
      return 0;
    cached.store(ncu * per);
  }
  return cached.load();
}

template <class L>
hipError_t sparse_l(const DevParts &d, const LoopArgs &la, const ChunkArgs &c1,
                    const ChunkArgs &c2, const TileArgs &ta,
                    const typename L::Rec *rec, bool persistent, int waves,
                    int max_tiles, bool rest, hipStream_t st) {
  const PosRec *pos = (const PosRec *)d.pos;
  if (!persistent) {
    const int nb = max_tiles > 0 ? xcd_blocks(max_tiles) : 0;
    if (nb > 0)
      hipLaunchKernelGGL((k_sparse_tiles<L>), dim3(nb), dim3(64), 0, st, d, la,
                         c1, c2, rec, pos, ta);
    if (rest)
      hipLaunchKernelGGL((k_sparse_tiles_rest<L>), dim3(kRestBlocks), dim3(64),
                         0, st, d, la, c1, c2, rec, pos, ta, nb);
    return hipGetLastError();
  }
  if (waves <= 0) waves = sparse_resident_blocks<L>();
  if (waves <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_sparse_queue<L>), dim3(waves), dim3(64), 0, st, d, la,
                     c1, c2, rec, pos, ta);
  return hipGetLastError();
}

hipError_t launch_sparse(LoopKind kind, const DevParts &d, const LoopArgs &la,
                         const ChunkArgs &c1, const ChunkArgs &c2,
                         const TileArgs &ta, bool persistent, int waves,
                         int max_tiles, bool rest, hipStream_t st) {
  if (c1.C != kSpC1 || c2.C != kSpC2) return hipErrorInvalidValue;
  switch (kind) {
    case LoopKind::Density:
      return sparse_l<DensityLoop>(d, la, c1, c2, ta, d.rd, persistent, waves,
                                   max_tiles, rest, st);
    case LoopKind::Gradient:
      return sparse_l<GradientLoop>(d, la, c1, c2, ta, d.rg, persistent, waves,
                                    max_tiles, rest, st);
    case LoopKind::Force:
      return sparse_l<ForceLoop>(d, la, c1, c2, ta, d.rf, persistent, waves,
                                 max_tiles, rest, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_fill_f32(float *p, float v, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f32, dim3(nblk(n, 256)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

template <class L, bool SUB>
hipError_t stencil_any(const DevParts &d, const LoopArgs &la,
                       const ChunkArgs &ca, const typename L::Rec *rec,
                       hipStream_t st) {
  switch (ca.C) {
    case 8: stencil_c<L, SUB, 8>(d, la, ca, rec, st); break;
    case 16: stencil_c<L, SUB, 16>(d, la, ca, rec, st); break;
    case 32: stencil_c<L, SUB, 32>(d, la, ca, rec, st); break;
    case 64: stencil_c<L, SUB, 64>(d, la, ca, rec, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class L, bool SUB, int C>
void task_c(const DevParts &d, const LoopArgs &la, const ChunkArgs &ca,
            const typename L::Rec *rec, const TaskArgs &ta, int nb,
            hipStream_t st) {
  hipLaunchKernelGGL((k_loop_task<L, C, SUB>), dim3(nb), dim3(64), 0, st, d, la,
                     ca, rec, (const PosRec *)d.pos, ta);
}

template <class L, bool SUB>
hipError_t task_any(const DevParts &d, const LoopArgs &la, const ChunkArgs &ca,
                    const typename L::Rec *rec, const TaskArgs &ta,
                    hipStream_t st) {
  const int nb = ta.ntiles[0] + ta.ntiles[1];
  if (nb <= 0) return hipSuccess;
  switch (ca.C) {
    case 8: task_c<L, SUB, 8>(d, la, ca, rec, ta, nb, st); break;
    case 16: task_c<L, SUB, 16>(d, la, ca, rec, ta, nb, st); break;
    case 32: task_c<L, SUB, 32>(d, la, ca, rec, ta, nb, st); break;
    case 64: task_c<L, SUB, 64>(d, la, ca, rec, ta, nb, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_stencil(LoopKind kind, bool subset, const DevParts &d,
                          const LoopArgs &la, const ChunkArgs &ca,
                          hipStream_t st) {
  if (ca.nchunks <= 0) return hipSuccess;
  switch (kind) {
    case LoopKind::Density:
      return subset ? stencil_any<DensityLoop, true>(d, la, ca, d.rd, st)
                    : stencil_any<DensityLoop, false>(d, la, ca, d.rd, st);
    case LoopKind::Gradient:
      return subset ? stencil_any<GradientLoop, true>(d, la, ca, d.rg, st)
                    : stencil_any<GradientLoop, false>(d, la, ca, d.rg, st);
    case LoopKind::Force:
      return stencil_any<ForceLoop, false>(d, la, ca, d.rf, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_task(LoopKind kind, bool subset, const DevParts &d,
                       const LoopArgs &la, const ChunkArgs &ca,
                       const TaskArgs &ta, hipStream_t st) {
  switch (kind) {
    case LoopKind::Density:
      return subset ? task_any<DensityLoop, true>(d, la, ca, d.rd, ta, st)
                    : task_any<DensityLoop, false>(d, la, ca, d.rd, ta, st);
    case LoopKind::Gradient:
      return task_any<GradientLoop, false>(d, la, ca, d.rg, ta, st);
    case LoopKind::Force:
      return task_any<ForceLoop, false>(d, la, ca, d.rf, ta, st);
  }
  return hipErrorInvalidValue;
}

template <class L>
hipError_t bundle_any(const DevParts &d, const LoopArgs &la,
                      const ChunkArgs &ca, const typename L::Rec *rec,
                      const BundleArgs &ba, hipStream_t st) {
  if (ba.nitems <= 0) return hipSuccess;
  const dim3 g(ba.nitems), b(64);
  const PosRec *p = (const PosRec *)d.pos;
  switch (ca.C) {
    case 8: hipLaunchKernelGGL((k_loop_bundle<L, 8>), g, b, 0, st, d, la, ca, rec, p, ba); break;
    case 16: hipLaunchKernelGGL((k_loop_bundle<L, 16>), g, b, 0, st, d, la, ca, rec, p, ba); break;
    case 32: hipLaunchKernelGGL((k_loop_bundle<L, 32>), g, b, 0, st, d, la, ca, rec, p, ba); break;
    case 64: hipLaunchKernelGGL((k_loop_bundle<L, 64>), g, b, 0, st, d, la, ca, rec, p, ba); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_bundle(LoopKind kind, const DevParts &d, const LoopArgs &la,
                         const ChunkArgs &ca, const BundleArgs &ba,
                         hipStream_t st) {
  switch (kind) {
    case LoopKind::Density:
      return bundle_any<DensityLoop>(d, la, ca, d.rd, ba, st);
    case LoopKind::Gradient:
      return bundle_any<GradientLoop>(d, la, ca, d.rg, ba, st);
    case LoopKind::Force:
      return bundle_any<ForceLoop>(d, la, ca, d.rf, ba, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_zero_u64(unsigned long long *p, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3(nblk(n, 64)), dim3(64), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_cell_mask(unsigned char *mask, int ncells, const int *cells,
                            int n, hipStream_t st) {
  if (ncells > 0)
    hipLaunchKernelGGL(k_zero_u8, dim3(nblk(ncells, 256)), dim3(256), 0, st,
                       mask, ncells);
  if (n > 0)
    hipLaunchKernelGGL(k_set_mask, dim3(nblk(n, 256)), dim3(256), 0, st, mask,
                       cells, n);
  return hipGetLastError();
}

hipError_t launch_cell_active(const DevParts &d, int n, int ncells,
                              int max_active_bin, unsigned char *act,
                              hipStream_t st) {
  if (ncells > 0)
    hipLaunchKernelGGL(k_zero_u8, dim3(nblk(ncells, 256)), dim3(256), 0, st,
                       act, ncells);
  if (n > 0)
    hipLaunchKernelGGL(k_cell_active, dim3(run_blocks(n, 256)), dim3(256), 0,
                       st, d, n, max_active_bin, act);
  return hipGetLastError();
}

hipError_t launch_invert_perm(const int *perm, int *inv, int n,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_invert_perm, dim3(nblk(n, 256)), dim3(256), 0, st, perm,
                     inv, n);
  return hipGetLastError();
}

hipError_t launch_mark_subset(int *flag, const int *inv, long long first,
                              const int *ind, int count, int lo, int hi,
                              unsigned int *err, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, (int *)err, 1);
  if (hi > lo)
    hipLaunchKernelGGL(k_zero_int, dim3(nblk(hi - lo, 256)), dim3(256), 0, st,
                       flag + lo, hi - lo);
  if (count > 0)
    hipLaunchKernelGGL(k_mark_subset, dim3(nblk(count, 256)), dim3(256), 0, st,
                       flag, inv, first, ind, count, lo, hi, err);
  return hipGetLastError();
}

hipError_t launch_share_flags(int *dst, const int *src, int lo, int hi,
                              int off, int nt, hipStream_t st) {
  if (hi <= lo) return hipSuccess;
  hipLaunchKernelGGL(k_share_flags, dim3(nblk(hi - lo, 256)), dim3(256), 0, st,
                     dst, src, lo, hi, off, nt);
  return hipGetLastError();
}

hipError_t launch_perm_moved(const int *perm, int n, unsigned int *moved,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, (int *)moved, 1);
  if (n > 0)
    hipLaunchKernelGGL(k_perm_moved, dim3(nblk(n, 256)), dim3(256), 0, st,
                       perm, n, moved);
  return hipGetLastError();
}

hipError_t launch_chunk_fill(const int *cell_start, const int *cell_chunk,
                             int ncells, int C, int nchunks, int n, int *start,
                             int *chunk_cell, hipStream_t st) {
  hipLaunchKernelGGL(k_chunk_fill, dim3(nblk(nchunks + 1, 256)), dim3(256), 0,
                     st, cell_start, cell_chunk, ncells, C, nchunks, n, start,
                     chunk_cell);
  return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned int *p, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_int, dim3(nblk(n, 64)), dim3(64), 0, st, (int *)p,
                     n);
  return hipGetLastError();
}

hipError_t launch_build_pos(const DevParts &d, const GridDesc &g, float pos_eps,
                            int lo, int hi, hipStream_t st) {
  if (hi <= lo) return hipSuccess;
  hipLaunchKernelGGL(k_build_pos, dim3(nblk(hi - lo, 256)), dim3(256), 0, st, d,
                     g, pos_eps, lo, hi);
  return hipGetLastError();
}

hipError_t launch_chunk_bounds(const DevParts &d, const int *chunk_start,
                               int c0, int c1, int C, double *bbox,
                               float *hg2max, unsigned int *hg2_all,
                               hipStream_t st, const GridDesc &g, float pos_eps,
                               float4 *fbox) {
  if (c1 <= c0) return hipSuccess;
  if (C != 8 && C != 16 && C != 32 && C != 64) return hipErrorInvalidValue;
  const int waves = nblk(c1 - c0, 64 / C);
  hipLaunchKernelGGL(k_chunk_bounds, dim3(nblk(waves, kBoundsBlock / 64)),
                     dim3(kBoundsBlock), 0, st, d, chunk_start, c0, c1, C, bbox,
                     hg2max, hg2_all, g, pos_eps, fbox);
  return hipGetLastError();
}

hipError_t launch_radius_bounds(const DevParts &d, const int *chunk_start,
                                int c0, int c1, int C, float *hg2max,
                                unsigned int *hg2_all, float pos_eps,
                                float4 *fbox, bool records, hipStream_t st) {
  if (c1 <= c0) return hipSuccess;
  if (C != 8 && C != 16 && C != 32 && C != 64) return hipErrorInvalidValue;
  const int waves = nblk(c1 - c0, 64 / C);
  hipLaunchKernelGGL(k_radius_bounds, dim3(nblk(waves, kBoundsBlock / 64)),
                     dim3(kBoundsBlock), 0, st, d, chunk_start, c0, c1, C,
                     hg2max, hg2_all, pos_eps, fbox,
                     records ? (PosRec *)d.pos : nullptr);
  return hipGetLastError();
}

hipError_t launch_build_records(LoopKind kind, const DevParts &d, int lo,
                                int hi, hipStream_t st, int max_active_bin,
                                const int *list, const int *cnt) {
  if (max_active_bin < 0) list = cnt = nullptr;  // all records: staged path
  if (hi <= lo) return hipSuccess;
  const int m = max_active_bin;
  // active records: kRuns runs per block; all records: one run, staged
  const dim3 g(m >= 0 ? run_blocks(hi - lo, kRecBlock) : nblk(hi - lo, kRecBlock)),
      b(kRecBlock);
  switch (kind) {
    case LoopKind::Density:
      hipLaunchKernelGGL(k_build_dens_rec, g, b, 0, st, d, lo, hi, m, list, cnt);
      break;
    case LoopKind::Gradient:
      hipLaunchKernelGGL(k_build_grad_rec, g, b, 0, st, d, lo, hi, m, list, cnt);
      break;
    case LoopKind::Force:
      hipLaunchKernelGGL(k_build_force_rec, g, b, 0, st, d, lo, hi, m, list, cnt);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_init_density(const DevParts &d, int lo, int hi,
                               int max_active_bin, float h_max, hipStream_t st,
                               const int *list, const int *cnt, bool light) {
  if (hi <= lo) return hipSuccess;
  hipLaunchKernelGGL(k_init_density, dim3(run_blocks(hi - lo, 256)), dim3(256), 0, st,
                     d, lo, hi, max_active_bin, h_max, list, cnt, light ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_ghost(const DevParts &d, const GhostArgs &ga, int lo, int hi,
                        unsigned long long *counters, hipStream_t st,
                        const int *list, const int *cnt, int off, int nt) {
  // counters[3] accumulates over the rounds of a pass (zeroed by the pass)
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, counters, 3);
  if (hi > lo)
    hipLaunchKernelGGL(k_ghost, dim3(run_blocks(hi - lo, kGhostBlock)),
                       dim3(kGhostBlock), 0, st, d, ga, lo, hi, off, nt,
                       counters, list, cnt);
  return hipGetLastError();
}

hipError_t launch_extra_ghost(const DevParts &d, const GhostArgs &ga, int lo,
                              int hi, hipStream_t st, const int *list,
                              const int *cnt) {
  if (hi <= lo) return hipSuccess;
  hipLaunchKernelGGL(k_extra_ghost, dim3(run_blocks(hi - lo, 256)), dim3(256), 0, st,
                     d, ga, lo, hi, list, cnt);
  return hipGetLastError();
}

hipError_t launch_end_force(const DevParts &d, int lo, int hi,
                            int max_active_bin, hipStream_t st, const int *list,
                            const int *cnt) {
  if (hi <= lo) return hipSuccess;
  hipLaunchKernelGGL(k_end_force, dim3(run_blocks(hi - lo, 256)), dim3(256), 0, st, d,
                     lo, hi, max_active_bin, list, cnt);
  return hipGetLastError();
}

hipError_t launch_copy_set(const CopySet &cs, int n, hipStream_t st) {
  if (n <= 0 || cs.m <= 0) return hipSuccess;
  if (cs.m > kMaxCopies) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_copy_set, dim3(nblk(n, 1024), cs.m), dim3(256), 0, st,
                     cs, n);
  return hipGetLastError();
}

hipError_t launch_count_active(const DevParts &d, int n, int max_active_bin,
                               unsigned int *out, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, (int *)out, 1);
  if (n > 0)
    hipLaunchKernelGGL(k_count_active, dim3(min(1024, run_blocks(n, 256))),
                       dim3(256), 0, st, (const int8_t *)d.tb, n, max_active_bin,
                       out);
  return hipGetLastError();
}

hipError_t launch_count(const DevParts &d, int n, int max_active_bin,
                        unsigned long long *out3, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, out3, 3);
  const int nb = n > 0 ? min(1024, run_blocks(n, 256)) : 1;
  hipLaunchKernelGGL(k_count, dim3(nb), dim3(256), 0, st, d, n, max_active_bin,
                     out3);
  return hipGetLastError();
}

}  // namespace sphk
