// sph_gather.h -- the neighbour-loop engine shared by the per-task and the
// batched kernels.
//
// One wave64 owns a tile of <= 64 consecutive particles of one cell, one
// particle i per lane (gather form of DOPAIR1/DOSELF1/DOPAIR2/DOSELF2,
// src/runner_doiact_functions_hydro.h:1234, 2299, 1601, 2624).  Candidates j
// stream through LDS in chunks of 64 records (one coalesced 16-byte-vector
// load per lane).  Test phase: all lanes run the FP32 pair test against the
// same j (broadcast LDS read) and set one bit of a per-lane 64-bit hit mask.
// Interaction phase: once a window of WIN chunks is tested, every lane walks
// its own set bits, reading its own j record from LDS.  Neighbour cells are
// visited in mirror pairs (offset o, then 26 - o) so that the lanes near one
// face and the lanes near the opposite face share a window and the
// interaction phase stays dense.  No atomics: accumulators live in
// registers and are written once.
#pragma once

#include "sph_loops.h"

namespace sphl {

// Visiting order of the 27 neighbour offsets o = 9(ox+1) + 3(oy+1) + (oz+1):
// the cell itself, then mirror pairs.
__constant__ const int kStencilOrder[27] = {13, 0,  26, 1,  25, 2,  24, 3,  23,
                                            4,  22, 5,  21, 6,  20, 7,  19, 8,
                                            18, 9,  17, 10, 16, 11, 15, 12, 14};

// What the LDS stage holds for a loop whose records are too large to gather
// from LDS: the position and search radius of j.
struct __attribute__((aligned(16))) PosRec {
  double x, y, z;
  float hg2, pad;
};

template <class L, bool FULL = L::kStageFull>
struct Stage {
  using T = typename L::Rec;
  static __device__ __forceinline__ void put(T &dst, const typename L::Rec &r) {
    dst = r;
  }
};
template <class L>
struct Stage<L, false> {
  using T = PosRec;
  static __device__ __forceinline__ void put(T &dst, const typename L::Rec &r) {
    dst.x = r.x;
    dst.y = r.y;
    dst.z = r.z;
    dst.hg2 = r.hg2;
    dst.pad = 0.f;
  }
};

template <class L, int WIN>
struct Window {
  unsigned long long m0, m1;  // hit masks of chunk slot 0 / 1
  int j0a, j0b;               // first candidate index of chunk slot 0 / 1
  int n;                      // chunks staged (uniform)
  double tx, ty, tz;          // image translation shared by the window
  double xs, ys, zs;          // x_i - t (per lane)
};

// Interaction phase over the window's hits.  With a full stage the records
// are read from buf[0, 64 * n); otherwise they are gathered from HBM
// (L1/L2: the stage just touched them).
template <class L, int WIN>
__device__ __forceinline__ void flush(const typename L::Rec *__restrict__ rec,
                                      const typename Stage<L>::T *buf,
                                      const typename L::I &s,
                                      Window<L, WIN> &w, typename L::Acc &acc) {
  unsigned long long m0 = w.m0, m1 = w.m1;
  for (;;) {
    int slot, jidx;
    if (m0) {
      slot = __builtin_ctzll(m0);
      jidx = w.j0a + slot;
      m0 &= m0 - 1;
    } else if (WIN > 1 && m1) {
      const int k = __builtin_ctzll(m1);
      slot = 64 + k;
      jidx = w.j0b + k;
      m1 &= m1 - 1;
    } else {
      break;
    }
    const typename L::Rec &j = [&]() -> const typename L::Rec & {
      if constexpr (L::kStageFull)
        return buf[slot];
      else
        return rec[jidx];
    }();
    float dx, dy, dz;
    const float r2 = pair_r2(w.xs, w.ys, w.zs, j.x, j.y, j.z, dx, dy, dz);
    L::interact(r2, dx, dy, dz, s, j, acc);
  }
  w.m0 = w.m1 = 0ull;
  w.n = 0;
}

// Stage one chunk of cnt <= 64 records into slot w.n and test it.  `self` is
// the lane's own particle index (excluded) or -1.
template <class L, int WIN>
__device__ __forceinline__ void add_chunk(
    const typename L::Rec *__restrict__ rec, typename Stage<L>::T *buf,
    const typename L::I &s, bool act, int self, int j0, int cnt,
    Window<L, WIN> &w, typename L::Acc &acc) {
  const int lane = threadIdx.x;
  typename Stage<L>::T *slot = buf + 64 * w.n;
  __syncthreads();
  if (lane < cnt) Stage<L>::put(slot[lane], rec[j0 + lane]);
  __syncthreads();
  unsigned int lo = 0u, hi = 0u;
  const int n_lo = cnt < 32 ? cnt : 32;
#pragma unroll 4
  for (int k = 0; k < n_lo; k++) {
    const typename Stage<L>::T &j = slot[k];
    float dx, dy, dz;
    const float r2 = pair_r2(w.xs, w.ys, w.zs, j.x, j.y, j.z, dx, dy, dz);
    lo |= L::test(r2, s, &j) ? (1u << k) : 0u;
  }
#pragma unroll 4
  for (int k = 32; k < cnt; k++) {
    const typename Stage<L>::T &j = slot[k];
    float dx, dy, dz;
    const float r2 = pair_r2(w.xs, w.ys, w.zs, j.x, j.y, j.z, dx, dy, dz);
    hi |= L::test(r2, s, &j) ? (1u << (k - 32)) : 0u;
  }
  unsigned long long m = ((unsigned long long)hi << 32) | lo;
  const int sk = self - j0;
  if (sk >= 0 && sk < 64) m &= ~(1ull << sk);
  if (!act) m = 0ull;
  if (w.n == 0) {
    w.m0 = m;
    w.j0a = j0;
  } else {
    w.m1 = m;
    w.j0b = j0;
  }
  w.n++;
}

// Adds the particles [j0, j0 + nj) of one neighbour cell whose image is
// translated by t.
template <class L, int WIN>
__device__ __forceinline__ void add_segment(
    const typename L::Rec *__restrict__ rec, typename Stage<L>::T *buf,
    const typename L::I &s, bool act, int self, int j0, int nj, double tx,
    double ty, double tz, Window<L, WIN> &w, typename L::Acc &acc) {
  if (nj <= 0) return;
  if (w.n > 0 && (tx != w.tx || ty != w.ty || tz != w.tz))
    flush<L, WIN>(rec, buf, s, w, acc);
  if (w.n == 0) {
    w.tx = tx;
    w.ty = ty;
    w.tz = tz;
    w.xs = s.x - tx;
    w.ys = s.y - ty;
    w.zs = s.z - tz;
  }
  for (int c = 0; c < nj; c += 64) {
    add_chunk<L, WIN>(rec, buf, s, act, self, j0 + c, min(64, nj - c), w, acc);
    if (w.n == WIN) flush<L, WIN>(rec, buf, s, w, acc);
  }
}

template <class L, int WIN>
__device__ __forceinline__ void window_init(Window<L, WIN> &w) {
  w.m0 = w.m1 = 0ull;
  w.j0a = w.j0b = 0;
  w.n = 0;
  w.tx = w.ty = w.tz = 0.0;
  w.xs = w.ys = w.zs = 0.0;
}

__device__ __forceinline__ void cell_coords(int cid, const GridDesc &g,
                                            int &cx, int &cy, int &cz) {
  cz = cid % g.nc[2];
  cy = (cid / g.nc[2]) % g.nc[1];
  cx = cid / (g.nc[2] * g.nc[1]);
}

// Neighbour o (0..26) of cell (cx,cy,cz) with the translation t of its image.
__device__ __forceinline__ int neighbour(const GridDesc &g, int cx, int cy,
                                         int cz, int o, double &tx, double &ty,
                                         double &tz) {
  int ox = o / 9 - 1, oy = (o / 3) % 3 - 1, oz = o % 3 - 1;
  int v[3] = {cx + ox, cy + oy, cz + oz};
  double t[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 3; k++) {
    if (v[k] < 0) {
      v[k] += g.nc[k];
      t[k] = -g.dim[k];
    } else if (v[k] >= g.nc[k]) {
      v[k] -= g.nc[k];
      t[k] = g.dim[k];
    }
  }
  tx = t[0];
  ty = t[1];
  tz = t[2];
  return (v[0] * g.nc[1] + v[1]) * g.nc[2] + v[2];
}

// Map a block to a tile so that each XCD (blocks are dealt round-robin over
// the 8 XCDs) works on a contiguous, spatially coherent range of tiles and
// reuses neighbour cells from its own L2.
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  const int per = (ntiles + 7) >> 3;
  return (b & 7) * per + (b >> 3);
}

template <class L, bool SUBSET>
__device__ __forceinline__ bool tile_begin(const DevParts &d, int max_active_bin,
                                           int first, int count,
                                           typename L::I &s,
                                           typename L::Acc &acc, bool &act) {
  const int lane = threadIdx.x;
  const int i = first + lane;
  act = lane < count;
  if (act) {
    act = is_active(d.tb[i], max_active_bin);
    if (SUBSET) act = act && d.redo[i];
  }
  if (!__any(act)) return false;  // the whole wave skips (uniform)
  if (act) {
    L::load_i(d, i, s);
    L::load_acc(d, i, acc);
  } else {
    // Lanes without an active particle run the tests masked out.
    s = typename L::I{};
    s.x = s.y = s.z = kFar;
  }
  return true;
}

// Batched: every tile of the tile list against its 27-cell stencil.
template <class L, bool SUBSET, int WIN>
__global__ __launch_bounds__(64) void k_loop_stencil(
    DevParts d, LoopArgs la, const typename L::Rec *__restrict__ rec,
    const int4 *__restrict__ tiles, int ntiles) {
  __shared__ typename Stage<L>::T buf[64 * WIN];
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  const int4 tl = tiles[t];
  typename L::I s;
  typename L::Acc acc;
  bool act;
  if (!tile_begin<L, SUBSET>(d, la.max_active_bin, tl.y, tl.z, s, acc, act))
    return;
  const int i = tl.y + threadIdx.x;
  Window<L, WIN> w;
  window_init(w);
  int cx, cy, cz;
  cell_coords(tl.x, la.grid, cx, cy, cz);
  for (int q = 0; q < 27; q++) {
    const int o = kStencilOrder[q];
    double tx, ty, tz;
    const int cj = neighbour(la.grid, cx, cy, cz, o, tx, ty, tz);
    const int j0 = la.cell_start[cj];
    const int nj = la.cell_start[cj + 1] - j0;
    add_segment<L, WIN>(rec, buf, s, act, o == 13 ? i : -1, j0, nj, tx, ty, tz,
                        w, acc);
  }
  if (w.n > 0) flush<L, WIN>(rec, buf, s, w, acc);
  if (act) L::store_acc(d, i, acc);
}

// Per task: up to two directed (i-cell <- j-cell) interactions; blocks
// [0, nt0) take tiles of the first i-cell, the rest the second.
template <class L, bool SUBSET, int WIN>
__global__ __launch_bounds__(64) void k_loop_task(
    DevParts d, LoopArgs la, const typename L::Rec *__restrict__ rec,
    TaskArgs ta) {
  __shared__ typename Stage<L>::T buf[64 * WIN];
  int q = 0;
  int b = blockIdx.x;
  if (b >= ta.ntiles[0]) {
    q = 1;
    b -= ta.ntiles[0];
  }
  const int ci = ta.icell[q];
  const int first = la.cell_start[ci] + 64 * b;
  const int count = min(64, la.cell_start[ci + 1] - first);
  typename L::I s;
  typename L::Acc acc;
  bool act;
  if (!tile_begin<L, SUBSET>(d, la.max_active_bin, first, count, s, acc, act))
    return;
  const int i = first + threadIdx.x;
  const int cj = ta.jcell[q];
  const int j0 = la.cell_start[cj];
  const int nj = la.cell_start[cj + 1] - j0;
  Window<L, WIN> w;
  window_init(w);
  add_segment<L, WIN>(rec, buf, s, act, ci == cj ? i : -1, j0, nj, ta.t[3 * q],
                      ta.t[3 * q + 1], ta.t[3 * q + 2], w, acc);
  if (w.n > 0) flush<L, WIN>(rec, buf, s, w, acc);
  if (act) L::store_acc(d, i, acc);
}

}  // namespace sphl
