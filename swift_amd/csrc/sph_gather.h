// sph_gather.h -- the neighbour-loop engine shared by the per-task and the
// batched kernels (gather form of DOPAIR1/DOSELF1/DOPAIR2/DOSELF2,
// src/runner_doiact_functions_hydro.h:1234, 2299, 1601, 2624).
//
// Particles are sorted by top-level cell and, inside a cell, along a Morton
// curve, then cut into chunks of C consecutive particles (C = 8 ... 64), so a
// chunk is spatially compact whatever the local density.  Every chunk has a
// bounding box and the largest search radius of its particles.
//
// One wave64 owns one i-chunk: lane = g + C * s holds particle g of the
// chunk and is "stream" s of S = 64 / C streams.  For each of the 27
// neighbour cells the wave tests the cell's chunks against the i-chunk box
// (one chunk per lane, then a ballot), and walks the surviving j-chunks in
// rows of S particles: stream s tests candidate (row start + s) against its
// particle g.  A hit sets one bit of the lane's row mask; the row table in
// LDS remembers where each row starts and which image translation it used.
// When 128 rows are tested, every lane walks its own set bits, gathers its
// j record and runs the interaction body, so the expensive bodies run dense
// although only a few per cent of the candidates are neighbours.  At the end
// the S partial accumulators of a particle are combined with cross-lane
// shuffles and written once.  No atomics: every directed interaction is
// evaluated by one lane.
#pragma once

#include "sph_loops.h"

namespace sphl {

// Visiting order of the 27 neighbour offsets o = 9(ox+1) + 3(oy+1) + (oz+1).
__constant__ const int kStencilOrder[27] = {13, 0,  26, 1,  25, 2,  24, 3,  23,
                                            4,  22, 5,  21, 6,  20, 7,  19, 8,
                                            18, 9,  17, 10, 16, 11, 15, 12, 14};

constexpr int kRows = 128;  // rows tested between interaction phases

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

// Map a block to a chunk so that each XCD (blocks are dealt round-robin over
// the 8 XCDs) works on a contiguous, spatially coherent range of chunks and
// reuses neighbour data from its own L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int per = (n + 7) >> 3;
  return (b & 7) * per + (b >> 3);
}

template <int S>
__device__ __forceinline__ float reduce_max_streams(float v) {
  // lanes g, g + C, g + 2C, ... hold the S streams of particle g
#pragma unroll
  for (int off = 64 / S; off < 64; off <<= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}

// Row state of the wave: 4 x 32-bit hit masks per lane, the row counter,
// and the row table in registers: lane k of `ra` (`rb`) describes row k
// (64 + k) as (neighbour slot q << 27) | first particle of the row.
struct Rows {
  unsigned int m0, m1, m2, m3;
  unsigned int ra, rb;
  int n;  // rows recorded (uniform)
};

constexpr unsigned int kRowMask = (1u << 27) - 1u;

// Interaction phase.  Called from converged code: the row table moves from
// registers to LDS first, because the per-lane loop below is divergent and a
// cross-lane read from a lane that has left the loop is not defined.
template <class L, int C>
__device__ __forceinline__ void interact_rows(
    const typename L::Rec *__restrict__ rec, const double *ttab,
    unsigned int *row_info, const typename L::I &st, int s, Rows &r,
    typename L::Acc &acc) {
  __syncthreads();
  row_info[threadIdx.x] = r.ra;
  row_info[64 + threadIdx.x] = r.rb;
  __syncthreads();
  unsigned int m0 = r.m0, m1 = r.m1, m2 = r.m2, m3 = r.m3;
  for (;;) {
    int row;
    if (m0) {
      row = __builtin_ctz(m0);
      m0 &= m0 - 1;
    } else if (m1) {
      row = 32 + __builtin_ctz(m1);
      m1 &= m1 - 1;
    } else if (m2) {
      row = 64 + __builtin_ctz(m2);
      m2 &= m2 - 1;
    } else if (m3) {
      row = 96 + __builtin_ctz(m3);
      m3 &= m3 - 1;
    } else {
      break;
    }
    const unsigned int info = row_info[row];
    const int j = (int)(info & kRowMask) + s;
    const int tc = (int)(info >> 27);
    const double xs = st.x - ttab[3 * tc];
    const double ys = st.y - ttab[3 * tc + 1];
    const double zs = st.z - ttab[3 * tc + 2];
    const typename L::Rec &jr = rec[j];
    float dx, dy, dz;
    const float r2 = pair_r2(xs, ys, zs, jr.x, jr.y, jr.z, dx, dy, dz);
    L::interact(r2, dx, dy, dz, st, jr, acc);
  }
  r.m0 = r.m1 = r.m2 = r.m3 = 0u;
  r.n = 0;
}

// One i-chunk against a list of neighbour cells.  `stencil` selects the 27
// neighbours of the i-chunk's cell; otherwise the single cell `cj0` with
// translation (tx0, ty0, tz0).
template <class L, int C, bool SUBSET>
__device__ __forceinline__ void run_chunk(
    const DevParts &d, const LoopArgs &la, const ChunkArgs &ca,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    int ich, bool stencil, int cj0, double tx0, double ty0, double tz0) {
  constexpr int S = 64 / C;           // candidate streams per particle
  constexpr int RPC = C / S > 0 ? C / S : 1;  // rows of a full chunk
  __shared__ PosRec lpos[64];
  __shared__ double ttab[27 * 3];
  __shared__ unsigned int row_info[kRows];
  const int lane = threadIdx.x;
  const int g = lane % C;
  const int s = lane / C;
  const int i0 = ca.chunk_start[ich], i1 = ca.chunk_start[ich + 1];
  const int i = i0 + g;
  bool act = i < i1;
  if (act) {
    act = is_active(d.tb[i], la.max_active_bin);
    if (SUBSET) act = act && d.redo[i];
  }
  if (!__any(act)) return;  // uniform
  typename L::I st;
  typename L::Acc acc;
  if (act) {
    L::load_i(d, i, st);
  } else {
    st = typename L::I{};
    st.x = st.y = st.z = kFar;
  }
  if (act && s == 0)
    L::load_acc(d, i, acc);
  else
    L::identity(acc);
  // i-chunk box and the largest search radius of its active particles
  const double *bb = ca.bbox + 6 * ich;
  const double ilx = bb[0], ily = bb[1], ilz = bb[2];
  const double ihx = bb[3], ihy = bb[4], ihz = bb[5];
  const float hg2i = wave_max(act ? st.hg2 : 0.f);

  Rows r;
  r.m0 = r.m1 = r.m2 = r.m3 = 0u;
  r.ra = r.rb = 0u;
  r.n = 0;
  int cx = 0, cy = 0, cz = 0;
  const int cell_i = ca.chunk_cell[ich];
  if (stencil) cell_coords(cell_i, la.grid, cx, cy, cz);
  const int nq = stencil ? 27 : 1;
  for (int q = 0; q < nq; q++) {
    double tx = tx0, ty = ty0, tz = tz0;
    int cj = cj0;
    if (stencil) cj = neighbour(la.grid, cx, cy, cz, kStencilOrder[q], tx, ty, tz);
    const int k0 = ca.cell_chunk[cj], k1 = ca.cell_chunk[cj + 1];
    if (k0 == k1) continue;
    const bool own = cj == cell_i;  // the lane's own particle is a candidate
    if (lane == 0) {
      ttab[3 * q] = tx;
      ttab[3 * q + 1] = ty;
      ttab[3 * q + 2] = tz;
    }
    const double xs = st.x - tx, ys = st.y - ty, zs = st.z - tz;
    const int jlo = ca.chunk_start[k0], jhi = ca.chunk_start[k1];
    for (int base = jlo; base < jhi; base += 64) {
      // The 64 / C chunks of this batch of <= 64 particles.
      const int kb0 = k0 + (base - jlo) / C;
      const int kb1 = min(k1, kb0 + 64 / C);
      const int kc = kb0 + lane;
      bool pass = false;
      if (kc < kb1) {
        const double *jb = ca.bbox + 6 * kc;
        const double gx = fmax(0.0, fmax(jb[0] + tx - ihx, ilx - (jb[3] + tx)));
        const double gy = fmax(0.0, fmax(jb[1] + ty - ihy, ily - (jb[4] + ty)));
        const double gz = fmax(0.0, fmax(jb[2] + tz - ihz, ilz - (jb[5] + tz)));
        const float gap2 = (float)(gx * gx + gy * gy + gz * gz);
        const float rad2 = L::kSymmetric ? fmaxf(hg2i, ca.hg2max[kc]) : hg2i;
        pass = gap2 <= rad2 * 1.0001f;
      }
      unsigned long long pm = __ballot(pass);
      if (pm == 0ull) continue;  // uniform
      // Stage the batch's candidate positions (one coalesced load per lane).
      const int cnt = min(64, jhi - base);
      __syncthreads();
      if (lane < cnt) lpos[lane] = pos[base + lane];
      __syncthreads();
      while (pm) {
        const int kk = kb0 + __builtin_ctzll(pm);
        pm &= pm - 1;
        const int j0 = ca.chunk_start[kk], j1 = ca.chunk_start[kk + 1];
#pragma unroll
        for (int rr = 0; rr < RPC; rr++) {
          const int rb = j0 + rr * S;
          if (rb >= j1) break;  // uniform
          const int j = rb + s;
          bool hit = false;
          if (rb + S <= j1 || j < j1) {
            const PosRec &p = lpos[j - base];
            float dx, dy, dz;
            const float r2 = pair_r2(xs, ys, zs, p.x, p.y, p.z, dx, dy, dz);
            hit = act && L::test(r2, st, &p);
            if (own) hit = hit && j != i;
          }
          const unsigned int bitv = hit ? (1u << (r.n & 31)) : 0u;
          const unsigned int info = ((unsigned int)q << 27) | (unsigned int)rb;
          switch (r.n >> 5) {
            case 0: r.m0 |= bitv; r.ra = lane == r.n ? info : r.ra; break;
            case 1: r.m1 |= bitv; r.ra = lane == r.n ? info : r.ra; break;
            case 2: r.m2 |= bitv; r.rb = lane == r.n - 64 ? info : r.rb; break;
            default: r.m3 |= bitv; r.rb = lane == r.n - 64 ? info : r.rb; break;
          }
          r.n++;
          if (r.n == kRows) interact_rows<L, C>(rec, ttab, row_info, st, s, r, acc);
        }
      }
    }
  }
  if (r.n > 0) {
    __syncthreads();
    interact_rows<L, C>(rec, ttab, row_info, st, s, r, acc);
  }
  L::template combine<S>(acc);
  if (act && s == 0) L::store_acc(d, i, acc);
}

// Batched: every chunk against its 27-cell stencil.
template <class L, int C, bool SUBSET>
__global__ __launch_bounds__(64) void k_loop_stencil(
    DevParts d, LoopArgs la, ChunkArgs ca,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos) {
  const int ich = xcd_remap(blockIdx.x, ca.nchunks);
  if (ich >= ca.nchunks) return;
  run_chunk<L, C, SUBSET>(d, la, ca, rec, pos, ich, true, 0, 0.0, 0.0, 0.0);
}

// Per task: up to two directed (i-cell <- j-cell) interactions; blocks
// [0, nt0) take the chunks of the first i-cell, the rest the second.
template <class L, int C, bool SUBSET>
__global__ __launch_bounds__(64) void k_loop_task(
    DevParts d, LoopArgs la, ChunkArgs ca,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    TaskArgs ta) {
  int q = 0;
  int b = blockIdx.x;
  if (b >= ta.ntiles[0]) {
    q = 1;
    b -= ta.ntiles[0];
  }
  const int ich = ca.cell_chunk[ta.icell[q]] + b;
  run_chunk<L, C, SUBSET>(d, la, ca, rec, pos, ich, false, ta.jcell[q],
                          ta.t[3 * q], ta.t[3 * q + 1], ta.t[3 * q + 2]);
}

}  // namespace sphl
