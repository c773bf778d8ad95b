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
constexpr int kSpHitCap = 96;          // hits per particle between walks
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
  double ttab[27 * 3];  // translation of each neighbour's image
  double qorg[27 * 3];  // corner of each neighbour's image
  double fo[27 * 6];    // pair-contract frame {oi, oj} of each neighbour
  int qk0[27], qk8[27], qpre[28], qj0[27], qj1[27];
  unsigned int plist[64];
  unsigned int pairs[kSpCI][kSpPairCap];
  unsigned int hits[kSpCI][kSpHitCap];
  float irad[kSpCI];  // squared radius of the per-particle box test
  float ipre[kSpCI];  // squared radius of the FP32 pre-test
  int ii[kSpCI];      // particle index
};

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
  if (act && s == 0 && !la.fresh)
    L::load_acc(d, i, acc);
  else
    L::identity(acc);
  // the particle's state: loaded by its stream-0 lane, kept in LDS
  double px = kFar, py = kFar, pz = kFar;
  float rt2 = 0.f;
  if (s == 0) {
    typename L::I st;
    if (act) {
      L::load_i(d, i, st);
    } else {
      st = typename L::I{};
      st.x = st.y = st.z = kFar;
    }
    sh.ist[g] = st;
    px = st.x;
    py = st.y;
    pz = st.z;
    rt2 = act ? st.hg2 : 0.f;
    sh.irad[g] = act ? rt2 : -1.f;
    sh.ipre[g] = act ? pretest_radius2(rt2, la.pos_eps) : -1.f;
    sh.ii[g] = act ? i : -1;
  }
  const float hg2i = wave_max(rt2);

  // Neighbour table: lane q < 27 describes stencil neighbour q.
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
    for (int k = 0; k < 3; k++) {
      sh.qorg[3 * lane + k] = (double)cj3[k] * w[k] + t[k];
      sh.ttab[3 * lane + k] = t[k];
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
  // the tile's box
  // (lanes 0..3 hold the particles; the reduction over them reaches lane 0)
  const bool pa = act && s == 0;
  double ilx = pa ? px : 1e300, ily = pa ? py : 1e300, ilz = pa ? pz : 1e300;
  double ihx = pa ? px : -1e300, ihy = pa ? py : -1e300,
         ihz = pa ? pz : -1e300;
#pragma unroll
  for (int off = 1; off < kSpCI; off <<= 1) {
    ilx = fmin(ilx, __shfl_xor(ilx, off));
    ily = fmin(ily, __shfl_xor(ily, off));
    ilz = fmin(ilz, __shfl_xor(ilz, off));
    ihx = fmax(ihx, __shfl_xor(ihx, off));
    ihy = fmax(ihy, __shfl_xor(ihy, off));
    ihz = fmax(ihz, __shfl_xor(ihz, off));
  }
  ilx = uniform(ilx);
  ily = uniform(ily);
  ilz = uniform(ilz);
  ihx = uniform(ihx);
  ihy = uniform(ihy);
  ihz = uniform(ihz);
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
      const double xg = uniform(sh.ist[gg].x), yg = uniform(sh.ist[gg].y),
                   zg = uniform(sh.ist[gg].z);
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
            const float xs = (float)(xg - sh.qorg[3 * q]);
            const float ys = (float)(yg - sh.qorg[3 * q + 1]);
            const float zs = (float)(zg - sh.qorg[3 * q + 2]);
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
      const double tx = sh.ttab[3 * q], ty = sh.ttab[3 * q + 1],
                   tz = sh.ttab[3 * q + 2];
      const double *jb = c1.bbox + 6 * kc;
      const double gx = fmax(0.0, fmax(jb[0] + tx - ihx, ilx - (jb[3] + tx)));
      const double gy = fmax(0.0, fmax(jb[1] + ty - ihy, ily - (jb[4] + ty)));
      const double gz = fmax(0.0, fmax(jb[2] + tz - ihz, ilz - (jb[5] + tz)));
      const float gap2 = (float)(gx * gx + gy * gy + gz * gz);
      const float rad2 = L::kSymmetric ? fmaxf(hg2i, c1.hg2max[kc]) : hg2i;
      pass = gap2 <= rad2 * 1.0001f;
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
        const float rad = sh.irad[g];
        if (sh.qj0[q] + js < sh.qj1[q] && rad >= 0.f) {
          const int k8 = sh.qk8[q] + kin * kSpSub + u;
          const double tx = sh.ttab[3 * q], ty = sh.ttab[3 * q + 1],
                       tz = sh.ttab[3 * q + 2];
          const double *jb = c2.bbox + 6 * k8;
          const double x = sh.ist[g].x, y = sh.ist[g].y, z = sh.ist[g].z;
          const double gx = fmax(0.0, fmax(jb[0] + tx - x, x - (jb[3] + tx)));
          const double gy = fmax(0.0, fmax(jb[1] + ty - y, y - (jb[4] + ty)));
          const double gz = fmax(0.0, fmax(jb[2] + tz - z, z - (jb[5] + tz)));
          const float gap2 = (float)(gx * gx + gy * gy + gz * gz);
          const float rad2 = L::kSymmetric ? fmaxf(rad, c2.hg2max[k8]) : rad;
          ok = gap2 <= rad2 * 1.0001f;
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
      // a plain read first: an exhausted queue costs no atomic (the tail
      // of a drain, and every wave of a grid for a nearly empty queue)
      base = __hip_atomic_load(&tq.heads[64 * q], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      if (q * per + base < hi) base = atomicAdd(&tq.heads[64 * q], kGrab);
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
