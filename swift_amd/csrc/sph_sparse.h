// sph_sparse.h -- neighbour loops for sparse tiles: a few active particles
// per cell, as on a clustered box where a small time bin is active.
//
// The chunk engine (sph_gather.h) tests every candidate of a chunk that
// reaches the tile's box against every particle of the tile.  With tiles of
// 4 scattered active particles in a dense cell, that box is as wide as a
// search radius and most of those tests fail: ~40 candidate tests per
// interaction, and the per-chunk bookkeeping is shared by 4 particles only.
// Here the candidates are culled per particle first:
//
//   1. chunks of 32 (the level-1 table) against the tile's box, one chunk
//      per lane, 64 at a time, survivors compacted into LDS;
//   2. the sub-chunks of 8 (the level-2 table: chunks of 8 of the same
//      cells, 4 per chunk of 32) of 4 surviving chunks at a time against
//      each particle of the tile separately (lane = chunk, sub-chunk,
//      particle), the passing (particle, sub-chunk) pairs going to one list
//      per particle;
//   3. the pairs of one particle at a time in rows of 8 pairs x 8
//      candidates: the FP32 pre-test of sph_gather.h (cell-relative
//      positions, radius widened by pretest_radius2), the hits appended to
//      the particle's hit list in LDS in row order.
//
// Steps 1 to 3 run in FP32 on the candidates' pre-test records (positions
// relative to their cell's corner): each chunk table carries the box of its
// chunks' records and their largest widened radius, and each particle its
// position relative to every neighbour image's corner, formed once per tile.
// A box gap formed with the same roundings is never larger than the
// pre-test distance of any member (float subtraction, products and FMAs of
// non-negative terms are monotone), so a box test with the pre-test's radii
// never rejects a candidate that the pre-test would pass.
//
// The 16 lanes of a particle then walk its hit list (lane s takes entries
// s, s + 16, ...), evaluate the pair contract in the FP64 frame of the
// hit's neighbour image and run the interaction body; the 16 partial sums
// are combined with shuffles and written once, as in the chunk engine.
// Every test widens, never narrows, the loop's own test, so the neighbour
// sets are those of the pair contract; only the order of summation differs
// from the chunk engine.
#pragma once

#include "sph_gather.h"

namespace sphl {

constexpr int kSpCI = kMinTile;        // particles per sparse tile
constexpr int kSpS = 64 / kSpCI;       // lanes per particle in the walk
constexpr int kSpC1 = kSparseChunk;    // level-1 chunk capacity
constexpr int kSpC2 = kSparseSub;      // sub-chunk capacity
constexpr int kSpSub = kSpC1 / kSpC2;  // sub-chunks per chunk
constexpr int kSpGroup = 64 / (kSpSub * kSpCI);  // chunks per pair test
constexpr int kSpPairCap = 64;         // pairs per particle between drains
constexpr int kSpHitCap = 80;          // hits per particle between walks
static_assert(kSpGroup * kSpSub * kSpCI == 64, "one pair test per lane");
// a row appends at most 64 hits to one list, a pair test at most
// kSpGroup * kSpSub pairs per particle (two of each per step)
static_assert(kSpHitCap >= 64, "a row's hits fit an empty list");
static_assert(kSpPairCap >= 2 * kSpGroup * kSpSub, "a step's pairs fit");

// LDS of one sparse wave.  Entries of the pair and hit lists are
// (neighbour q << 27) | (particle index - first particle of q's cell).
// The tile's particle states live here between walks (registers are kept
// for the culls and the accumulators).
template <class I>
struct SparseLds {
  I ist[kSpCI];
  double fo[27 * 6];    // pair-contract frame {oi, oj} of each neighbour
  // particle g relative to the corner of neighbour image q (FP32 pre-test
  // frame), entry 4 q + g; and the tile's box in it (active particles)
  float frx[27 * kSpCI], fry[27 * kSpCI], frz[27 * kSpCI];
  float tlo[27 * 3], thi[27 * 3];
  int qk0[27], qk8[27], qpre[28], qj0[27], qj1[27];
  unsigned int plist[64];
  unsigned int pairs[kSpCI][kSpPairCap];
  unsigned int hits[kSpCI][kSpHitCap];
  float ipre[kSpCI];  // squared radius of the FP32 pre-test (-1: inactive)
  int ii[kSpCI];      // record slot of each particle (-1: inactive)
};

// Squared gap between two FP32 boxes (0 when they overlap), with the
// roundings of the pre-test distance.
__device__ __forceinline__ float box_gap2(float alx, float aly, float alz,
                                          float ahx, float ahy, float ahz,
                                          float4 blo, float4 bhi) {
  const float gx = fmaxf(0.f, fmaxf(blo.x - ahx, alx - bhi.x));
  const float gy = fmaxf(0.f, fmaxf(blo.y - ahy, aly - bhi.y));
  const float gz = fmaxf(0.f, fmaxf(blo.z - ahz, alz - bhi.z));
  return fmaf(gz, gz, fmaf(gy, gy, gx * gx));
}

// Count field g (8 bits) of a packed per-particle counter.
__device__ __forceinline__ int pk_get(unsigned int p, int g) {
  return (int)((p >> (8 * g)) & 0xffu);
}

// Number of set bits of m in the lanes below this one.
__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi(
      (unsigned int)(m >> 32),
      __builtin_amdgcn_mbcnt_lo((unsigned int)m, 0u));
}

// One sparse tile of the work queue against the 27-cell stencil of its
// cell.  c1: the level-1 table (chunks of 32), c2: the level-2 table
// (chunks of 8 of the same cells).
template <class L>
__device__ __forceinline__ void run_sparse(
    const DevParts &d, const LoopArgs &la, const ChunkArgs &c1,
    const ChunkArgs &c2, const typename L::Rec *__restrict__ rec,
    const PosRec *__restrict__ pos, const TileArgs &tq, int tile) {
  constexpr int S = kSpS;
  using Rec = typename L::Rec;
  __shared__ SparseLds<typename L::I> sh;
  const int lane = threadIdx.x & 63;
  const int g = lane % kSpCI;
  const int s = lane / kSpCI;
  // the tile builder applied the loop's predicate (and filt narrows it)
  const int tl = tq.ids ? uniform(tq.ids[tile]) : tile;
  const int ts = uniform(tq.tile_start[tl]);
  bool act = g < uniform(tq.tile_start[tl + 1]) - ts;
  const int i = act ? tq.alist[ts + g] : 0;
  if (tq.filt && act) act = tq.filt[i] != 0;
  if (!__any(act)) return;  // uniform
  const int cell_i = uniform(tq.tile_cell[tl]);
  typename L::Acc acc;
  L::identity(acc);
  // the particle's state: loaded by its stream-0 lane, kept in LDS
  float pre = -1.f;
  if (s == 0) {
    typename L::I st;
    if (act) {
      L::load_i(d, i, st);
      if (la.fresh)
        L::fresh_acc(st, acc);
      else
        L::load_acc(d, i, acc);
    } else {
      st = typename L::I{};
      st.x = st.y = st.z = kFar;
    }
    sh.ist[g] = st;
    pre = act ? pretest_radius2(st.hg2, la.pos_eps) : -1.f;
    sh.ipre[g] = pre;
    sh.ii[g] = act ? rec_slot(d, i) : -1;  // the self pair, in record slots
  }
  const float maxpre = wave_max(pre);
  wave_sync();

  // Neighbour table: lane q < 27 describes stencil neighbour q, and forms
  // the particles' pre-test frames in its image and the tile's box there.
  const double w[3] = {la.grid.dim[0] / la.grid.nc[0],
                       la.grid.dim[1] / la.grid.nc[1],
                       la.grid.dim[2] / la.grid.nc[2]};
  int ci3[3];
  cell_coords(cell_i, la.grid, ci3[0], ci3[1], ci3[2]);
  int nchq = 0;
  bool qself = false;
  if (lane < 27) {
    const int o = kStencilOrder[lane];
    double t[3];
    const int cj = neighbour(la.grid, ci3[0], ci3[1], ci3[2], o, t[0], t[1],
                             t[2]);
    int cj3[3];
    cell_coords(cj, la.grid, cj3[0], cj3[1], cj3[2]);
    const int o3[3] = {o / 9 - 1, (o / 3) % 3 - 1, o % 3 - 1};
    double org[3];
    for (int k = 0; k < 3; k++) org[k] = (double)cj3[k] * w[k] + t[k];
    float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
    for (int gg = 0; gg < kSpCI; gg++) {
      const float fx = (float)(sh.ist[gg].x - org[0]);
      const float fy = (float)(sh.ist[gg].y - org[1]);
      const float fz = (float)(sh.ist[gg].z - org[2]);
      sh.frx[kSpCI * lane + gg] = fx;
      sh.fry[kSpCI * lane + gg] = fy;
      sh.frz[kSpCI * lane + gg] = fz;
      if (sh.ipre[gg] >= 0.f) {
        lo[0] = fminf(lo[0], fx);
        lo[1] = fminf(lo[1], fy);
        lo[2] = fminf(lo[2], fz);
        hi[0] = fmaxf(hi[0], fx);
        hi[1] = fmaxf(hi[1], fy);
        hi[2] = fmaxf(hi[2], fz);
      }
    }
    for (int k = 0; k < 3; k++) {
      sh.tlo[3 * lane + k] = lo[k];
      sh.thi[3 * lane + k] = hi[k];
    }
    double fo[6];
    qself = pair_frame(o3, ci3, cj3, t, w, fo);
    for (int k = 0; k < 6; k++) sh.fo[6 * lane + k] = fo[k];
    const int k0 = c1.cell_chunk[cj], k1 = c1.cell_chunk[cj + 1];
    sh.qk0[lane] = k0;
    sh.qk8[lane] = c2.cell_chunk[cj];
    sh.qj0[lane] = la.cell_start[cj];
    sh.qj1[lane] = la.cell_start[cj + 1];
    nchq = k1 - k0;
  }
  const unsigned int selfm =
      (unsigned int)uniform((int)(unsigned int)__ballot(qself));
  int incl = nchq;  // inclusive prefix over lanes 0..31
#pragma unroll
  for (int off = 1; off < 32; off <<= 1) {
    const int v = __shfl_up(incl, off);
    if (lane >= off) incl += v;
  }
  if (lane < 27) sh.qpre[lane] = incl - nchq;
  const int total = uniform(__shfl(incl, 26));
  wave_sync();

  // The walk over the hit lists (all lanes; hcp holds the list lengths).
  // The frame of the lane's current neighbour image is kept in registers:
  // dx = a - (float)(x_j - b), b = x_i and a = 0 in the i-cell itself,
  // b = oj and a = (float)(x_i - oi) across cells (pair_frame).
  auto walk = [&](unsigned int hcp) {
    const int n = pk_get(hcp, g);
    if (s >= n) return;
    const typename L::I st = sh.ist[g];
    int cq = -1, cj0 = 0;
    float ax = 0.f, ay = 0.f, az = 0.f;
    double bx = 0.0, by = 0.0, bz = 0.0;
    for (int k = s; k < n; k += S) {
      const unsigned int e = sh.hits[g][k];
      const int q = (int)(e >> 27);
      if (q != cq) {
        cq = q;
        cj0 = sh.qj0[q];
        const bool self = (selfm >> q) & 1u;
        const double *fo = &sh.fo[6 * q];
        bx = self ? st.x : fo[3];
        by = self ? st.y : fo[4];
        bz = self ? st.z : fo[5];
        ax = self ? 0.f : (float)(st.x - fo[0]);
        ay = self ? 0.f : (float)(st.y - fo[1]);
        az = self ? 0.f : (float)(st.z - fo[2]);
      }
      const int j = cj0 + (int)(e & kRowMask);
      const Rec r = rec[j];
      const float dx = ax - (float)(r.x - bx);
      const float dy = ay - (float)(r.y - by);
      const float dz = az - (float)(r.z - bz);
      const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
      if (L::test(r2, st, &r)) L::interact(r2, dx, dy, dz, st, r, acc);
    }
  };

  unsigned int pcp = 0u;  // pairs per particle (8 bits each)
  unsigned int hcp = 0u;  // hits per particle (8 bits each)
  // Step 3: the pairs of every particle in rows of 8 pairs x 8 candidates.
  auto drain = [&]() {
    wave_sync();
    for (int gg = 0; gg < kSpCI; gg++) {
      const int np = pk_get(pcp, gg);
      if (np == 0) continue;  // uniform
      const float pre = __int_as_float(
          __builtin_amdgcn_readfirstlane(__float_as_int(sh.ipre[gg])));
      const int ig = uniform(sh.ii[gg]);
      // pre-test of candidate lane % 8 of pair p
      auto row = [&](int p, bool &pass, unsigned int &ent) {
        pass = false;
        ent = 0u;
        if (p < np) {
          const unsigned int e = sh.pairs[gg][p];
          const int q = (int)(e >> 27);
          const int c0 = sh.qj0[q];
          const int j = c0 + (int)(e & kRowMask) + lane % kSpC2;
          if (j < sh.qj1[q]) {
            const PosRec pj = pos[j];
            const float xs = sh.frx[kSpCI * q + gg];
            const float ys = sh.fry[kSpCI * q + gg];
            const float zs = sh.frz[kSpCI * q + gg];
            const float dx = xs - pj.x, dy = ys - pj.y, dz = zs - pj.z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            pass = L::kSymmetric ? (r2 < pre || r2 < pj.hg2m) : r2 < pre;
            pass = pass && j != ig;
            ent = ((unsigned int)q << 27) | (unsigned int)(j - c0);
          }
        }
      };
      // two rows per step, their loads in flight together
      constexpr int kRow = 64 / kSpC2;  // pairs per row
      for (int pr = 0; pr < np; pr += 2 * kRow) {
        bool p0, p1;
        unsigned int e0, e1;
        row(pr + lane / kSpC2, p0, e0);
        row(pr + kRow + lane / kSpC2, p1, e1);
        const unsigned long long h0 = __ballot(p0), h1 = __ballot(p1);
        const int n0 = __popcll(h0), n1 = __popcll(h1);
        if (n0 + n1 == 0) continue;  // uniform
        // each row appends at most 64 hits: walk first when they do not fit
        if (pk_get(hcp, gg) + n0 > kSpHitCap) {  // uniform
          wave_sync();
          walk(hcp);
          wave_sync();
          hcp = 0u;
        }
        if (p0) sh.hits[gg][pk_get(hcp, gg) + lanes_below(h0)] = e0;
        hcp += (unsigned int)n0 << (8 * gg);
        if (pk_get(hcp, gg) + n1 > kSpHitCap) {  // uniform
          wave_sync();
          walk(hcp);
          wave_sync();
          hcp = 0u;
        }
        if (p1) sh.hits[gg][pk_get(hcp, gg) + lanes_below(h1)] = e1;
        hcp += (unsigned int)n1 << (8 * gg);
      }
    }
    wave_sync();
    pcp = 0u;
  };

  for (int fb = 0; fb < total; fb += 64) {
    // Step 1: one level-1 chunk per lane against the tile's box.
    const int f = fb + lane;
    bool pass = false;
    unsigned int pe = 0u;
    if (f < total) {
      int q = 0;
#pragma unroll
      for (int b = 16; b > 0; b >>= 1)
        if (q + b < 27 && sh.qpre[q + b] <= f) q += b;
      const int kc = sh.qk0[q] + (f - sh.qpre[q]);
      const float4 blo = c1.fbox[2 * kc], bhi = c1.fbox[2 * kc + 1];
      const float gap2 =
          box_gap2(sh.tlo[3 * q], sh.tlo[3 * q + 1], sh.tlo[3 * q + 2],
                   sh.thi[3 * q], sh.thi[3 * q + 1], sh.thi[3 * q + 2], blo,
                   bhi);
      pass = gap2 < (L::kSymmetric ? fmaxf(maxpre, blo.w) : maxpre);
      pe = ((unsigned int)q << 27) | (unsigned int)(kc - sh.qk0[q]);
    }
    const unsigned long long pm = __ballot(pass);
    if (pm == 0ull) continue;  // uniform
    const int nsurv = __popcll(pm);
    if (pass) sh.plist[lanes_below(pm)] = pe;
    wave_sync();
    // Step 2: sub-chunks of 2 kSpGroup surviving chunks at a time against
    // each particle (two tests per lane, their loads in flight together).
    auto sub_test = [&](int c, bool &ok, unsigned int &ent) {
      const int u = (lane / kSpCI) % kSpSub;
      ok = false;
      ent = 0u;
      if (c < nsurv) {
        const unsigned int e = sh.plist[c];
        const int q = (int)(e >> 27);
        const int kin = (int)(e & kRowMask);  // chunk of 32 within the cell
        const int js = kin * kSpC1 + u * kSpC2;  // first particle, in-cell
        const float pg = sh.ipre[g];
        if (sh.qj0[q] + js < sh.qj1[q] && pg >= 0.f) {
          const int k8 = sh.qk8[q] + kin * kSpSub + u;
          const float4 blo = c2.fbox[2 * k8], bhi = c2.fbox[2 * k8 + 1];
          const float x = sh.frx[kSpCI * q + g], y = sh.fry[kSpCI * q + g],
                      z = sh.frz[kSpCI * q + g];
          const float gap2 = box_gap2(x, y, z, x, y, z, blo, bhi);
          ok = gap2 < (L::kSymmetric ? fmaxf(pg, blo.w) : pg);
          ent = ((unsigned int)q << 27) | (unsigned int)js;
        }
      }
    };
    // lanes of particle g: every kSpCI-th lane from g
    const unsigned long long gmask = 0x1111111111111111ull << g;
    for (int pb = 0; pb < nsurv; pb += 2 * kSpGroup) {
      bool any_full = false;
#pragma unroll
      for (int gg = 0; gg < kSpCI; gg++)
        any_full = any_full ||
                   pk_get(pcp, gg) + 2 * kSpGroup * kSpSub > kSpPairCap;
      if (any_full) drain();  // uniform
      bool o0, o1;
      unsigned int e0, e1;
      sub_test(pb + lane / (kSpSub * kSpCI), o0, e0);
      sub_test(pb + kSpGroup + lane / (kSpSub * kSpCI), o1, e1);
      const unsigned long long m0 = __ballot(o0), m1 = __ballot(o1);
      const int pc = pk_get(pcp, g);
      if (o0) sh.pairs[g][pc + lanes_below(m0 & gmask)] = e0;
      if (o1)
        sh.pairs[g][pc + __popcll(m0 & gmask) + lanes_below(m1 & gmask)] = e1;
#pragma unroll
      for (int gg = 0; gg < kSpCI; gg++)
        pcp += (unsigned int)(__popcll(m0 & (0x1111111111111111ull << gg)) +
                              __popcll(m1 & (0x1111111111111111ull << gg)))
               << (8 * gg);
    }
    wave_sync();
  }
  drain();
  if (hcp != 0u) walk(hcp);  // uniform
  L::template combine<S>(acc);
  if (act && s == 0) L::store_acc(d, i, acc);
}

// Grid form: one wave per sparse tile (as k_loop_tiles).
template <class L>
__global__ __launch_bounds__(64, L::kSparseWavesPerSimd) void k_sparse_tiles(
    DevParts d, LoopArgs la, ChunkArgs c1, ChunkArgs c2,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    TileArgs tq) {
  const int t = xcd_item(blockIdx.x, threadIdx.x >> 6, *tq.ntiles);
  if (t >= *tq.ntiles) return;  // uniform per wave
  run_sparse<L>(d, la, c1, c2, rec, pos, tq, t);
}

// The tiles beyond a grid of b0 blocks (as k_loop_tiles_rest).
template <class L>
__global__ __launch_bounds__(64) void k_sparse_tiles_rest(
    DevParts d, LoopArgs la, ChunkArgs c1, ChunkArgs c2,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    TileArgs tq, int b0) {
  const int n = *tq.ntiles;
  const int nb = xcd_blocks(n);
  for (int b = b0 + (int)blockIdx.x; b < nb; b += (int)gridDim.x) {
    const int t = xcd_item(b, threadIdx.x >> 6, n);
    if (t >= n) continue;  // uniform per wave
    wave_sync();
    run_sparse<L>(d, la, c1, c2, rec, pos, tq, t);
  }
}

// Persistent form (as k_loop_queue): waves drain the eight per-XCD queues.
template <class L>
__global__ __launch_bounds__(64, kOccSparseQueue) void k_sparse_queue(
    DevParts d, LoopArgs la, ChunkArgs c1, ChunkArgs c2,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    TileArgs tq) {
  const int nt = *tq.ntiles;
  const int per = (nt + 7) >> 3;
  int q = blockIdx.x & 7;
  for (int tries = 0; tries < 8;) {
    const int hi = min(nt, (q + 1) * per);
    int base = 0;
    if ((threadIdx.x & 63) == 0) {
      // Once its own queue is exhausted, a wave reads the others' heads
      // before it tries to steal: an exhausted queue then costs no atomic
      // (the tail of a drain, and every wave of a grid for a nearly empty
      // queue); its own queue takes the atomic directly.
      base = tries > 0 ? __hip_atomic_load(&tq.heads[64 * q], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT)
                       : 0;
      if (tries == 0 || q * per + base < hi)
        base = atomicAdd(&tq.heads[64 * q], kGrab);
    }
    base = uniform(__shfl(base, 0));
    const int lo = q * per + base;
    if (lo >= hi) {  // uniform: this queue is exhausted
      q = (q + 1) & 7;
      tries++;
      continue;
    }
    const int end = min(hi, lo + kGrab);
    for (int t = lo; t < end; t++) {
      wave_sync();
      run_sparse<L>(d, la, c1, c2, rec, pos, tq, t);
    }
  }
}

}  // namespace sphl
