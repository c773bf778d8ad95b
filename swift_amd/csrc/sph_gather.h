// sph_gather.h -- the neighbour-loop engine shared by the per-task and the
// batched kernels (gather form of DOPAIR1/DOSELF1/DOPAIR2/DOSELF2,
// src/runner_doiact_functions_hydro.h:1234, 2299, 1601, 2624).
//
// Particles are sorted by top-level cell and, inside a cell, along a Morton
// curve, then cut into chunks of CJ consecutive particles (CJ = 8 ... 64), so
// a chunk is spatially compact whatever the local density.  Every chunk has
// a bounding box and the largest search radius of its particles.
//
// One wave64 owns one i-chunk or tile of CI particles: lane = g + CI * s
// holds particle g and is "stream" s of S = 64 / CI streams (CJ = CI for
// i-chunks taken from the chunk table; a work-queue tile may face candidate
// chunks of half its size).  The wave first builds its
// neighbour table (the 27 neighbour cells, their images and chunk ranges:
// one lane per cell, all loads in flight together), then culls the flat list
// of candidate chunks 64 at a time (one chunk per lane: box distance against
// the largest search radius, then a ballot) and compacts the survivors into
// LDS.  Survivors are staged 64 particles at a time into LDS and tested in
// rows of S: stream s tests candidate (row start + s) against its particle
// g.  The row test is an FP32 pre-test on cell-relative positions whose
// radius is widened by a proven bound on its difference from the pair
// contract (pretest_radius2), so it never rejects a neighbour.  A pass sets
// one bit of the lane's row masks (kMasks x 64 bits); a per-chunk table in
// LDS remembers where each chunk starts and which image it uses.  After
// 64 x kMasks rows every lane
// walks its own set bits, gathers its j record, evaluates the exact FP64
// contract and test, and runs the interaction body, so the expensive bodies
// run dense although only a few per cent of the candidates are neighbours.
// At the end the S partial accumulators of a particle are combined with
// cross-lane shuffles and written once.  No atomics: every directed
// interaction is evaluated by one lane.
//
// The BUILD form (density loop, C = kListChunk) tests against the skinned
// radius lam2 * (gamma_K h_i)^2 instead, appends every hit to the lane's
// stream of the particle's neighbour list, and interacts with the hits
// inside the plain radius.  Per stream, hits are visited in the same order
// as by the plain form, so both give bit-identical sums.
#pragma once

#include <type_traits>

#include "sph_loops.h"

namespace sphl {

// Visiting order of the 27 neighbour offsets o = 9(ox+1) + 3(oy+1) + (oz+1).
__constant__ const int kStencilOrder[27] = {13, 0,  26, 1,  25, 2,  24, 3,  23,
                                            4,  22, 5,  21, 6,  20, 7,  19, 8,
                                            18, 9,  17, 10, 16, 11, 15, 12, 14};

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

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}

constexpr unsigned int kRowMask = (1u << 27) - 1u;

// Two 64-bit hit masks per lane between two interaction phases: 128 rows
// per phase even out the lanes' hit counts better than 64 (-2 % per pass).
constexpr int kMasks = 2;

// The candidate pre-test of the density and gradient loops runs in packed
// FP16 (two rows per instruction) when a chunk has an even number of rows.
typedef _Float16 half2v __attribute__((ext_vector_type(2)));

// One wave per block in the batched kernels: each wave works on its own
// chunk or tile with its own LDS (several waves per block measured slower on
// the 128^3 box).
constexpr int kWaves = 1;

// Launch bounds of the batched neighbour kernels: one wave per block and
// the loop's occupancy target (its own one for the one-tile-per-wave
// kernel).
#define SWSPH_LOOP_BOUNDS(L) __launch_bounds__(64, L::kWavesPerSimd)
#define SWSPH_TILE_BOUNDS(L) __launch_bounds__(64, L::kTileWavesPerSimd)

// A value every lane of the wave holds (e.g. read from LDS at a common
// address), moved to a scalar register so that branches on it stay uniform.
__device__ __forceinline__ int uniform(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ double uniform(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)b);
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __hiloint2double(hi, lo);
}

// Orders LDS accesses of one wave around a point where lanes exchange data
// through LDS (a wave's LDS operations execute in issue order; this stops
// the compiler from moving them across).
__device__ __forceinline__ void wave_sync() {
  __syncthreads();  // the block is the wave
}

// Block b of a batched grid over n work items, wave w of it: each XCD (the
// blocks are dealt round-robin over the 8 XCDs) works on a contiguous,
// spatially coherent range of items and reuses neighbour data from its own
// L2.  Returns n for the waves past the end.
__device__ __forceinline__ int xcd_item(int b, int w, int n) {
  const int per = (n + 7) >> 3;
  const int k = (b >> 3) * kWaves + w;
  return k < per ? min(n, (b & 7) * per + k) : n;
}

// Blocks of a batched grid over n items (see xcd_item).
__host__ __device__ inline int xcd_blocks(int n) {
  const int per = (n + 7) >> 3;
  return 8 * ((per + kWaves - 1) / kWaves);
}

// LDS of one wave: the neighbour table, the compacted surviving chunks, the
// staged candidate positions and the chunk table of the current 64 rows.
// CI: i-particles per wave (S = 64 / CI candidate streams each); CJ:
// particles per candidate chunk (CJ >= S: a row of S candidates lies in one
// chunk).
template <int CI, int CJ>
struct GatherLds {
  static constexpr int S = 64 / CI;
  static_assert(CJ >= S && CJ <= 64, "a row must lie in one chunk");
  static constexpr int RPC = CJ / S;        // rows per chunk
  static constexpr int SLOTS = 64 * kMasks / RPC;  // chunks per phase
  PosRec lpos[64];
  // FP16 pre-test: staged candidates in the tile frame, rows paired
  _Float16 hx[64], hy[64], hz[64];
  float qoffc[27 * 3];  // neighbour image corners relative to the tile centre
  double ttab[27 * 3];
  double qorg[27 * 3];  // corner of each neighbour's image
  double fo[27 * 6];    // pair-contract frame {oi, oj} of each neighbour
  unsigned int ocode[27];  // stencil offset code of each neighbour (lists)
  int qk0[27], qpre[28], qj0[27], qj1[27], qcell[27];
  unsigned int plist[64];
  unsigned int cinfo[SLOTS];
};

// Interaction phase over the rows recorded in the lanes' masks.  Called from
// converged code; the per-lane loop below is divergent, so everything it
// shares is in LDS.  Each lane walks its own hits one at a time (a
// software-pipelined walk that holds the next record in a second register
// set spilled registers in the density loop and was slower).
template <class L, int CI, int CJ, bool BUILD>
__device__ __forceinline__ void interact_rows(
    const typename L::Rec *__restrict__ rec, const GatherLds<CI, CJ> &sh,
    const typename L::I &st, int s, unsigned long long m0,
    unsigned long long m1, typename L::Acc &acc, const NbrList &nl, int i,
    float rt2, int &ncnt, unsigned int selfm) {
  constexpr int S = GatherLds<CI, CJ>::S, RPC = GatherLds<CI, CJ>::RPC;
  // next row of this lane (ascending), taken out of the masks
  auto pop = [&]() -> int {
    const bool lo = m0 != 0ull;
    const unsigned long long mm = lo ? m0 : m1;
    const int row = (mm ? __builtin_ctzll(mm) : 0) + (lo ? 0 : 64);
    const unsigned long long rest = mm & (mm - 1ull);
    m0 = lo ? rest : m0;
    m1 = lo ? m1 : rest;
    return row;
  };
  auto any = [&]() { return (m0 | m1) != 0ull; };
  // record index and image of a row
  auto locate = [&](int row, int &q) -> int {
    const unsigned int info = sh.cinfo[row / RPC];
    q = (int)(info >> 27);
    return (int)(info & kRowMask) + (row % RPC) * S + s;
  };
  using Rec = typename L::Rec;
  // The pair contract in the frame of the lane's current neighbour image
  // cq: dx = a - (float)(x_j - b), with b = x_i and a = 0 in the i-cell
  // itself (the FP64 difference rounded once; (float)(x_i - x_j) equals
  // 0 - (float)(x_j - x_i) bit for bit) and b = oj, a = (float)(x_i - oi)
  // across cells (pair_frame).  Hits come in neighbour order, so the
  // frame changes a few times per walk and is kept in registers.
  int cq = -1;
  float ax = 0.f, ay = 0.f, az = 0.f;
  double bx = 0.0, by = 0.0, bz = 0.0;
  // Evaluates hit (j, q) with record r.
  auto eval = [&](int j, int q, const Rec &r) {
    if (q != cq) {
      cq = q;
      const bool self = (selfm >> q) & 1u;
      const double *fo = &sh.fo[6 * q];
      bx = self ? st.x : fo[3];
      by = self ? st.y : fo[4];
      bz = self ? st.z : fo[5];
      ax = self ? 0.f : (float)(st.x - fo[0]);
      ay = self ? 0.f : (float)(st.y - fo[1]);
      az = self ? 0.f : (float)(st.z - fo[2]);
    }
    const float dx = ax - (float)(r.x - bx);
    const float dy = ay - (float)(r.y - by);
    const float dz = az - (float)(r.z - bz);
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    if (BUILD) {
      if (r2 < rt2) {
        if (ncnt < nl.Ks)
          nl.idx[((size_t)ncnt * nl.npad + i) * kListStreams + s] =
              (sh.ocode[q] << 27) | (unsigned int)j;
        ncnt++;
        if (L::test(r2, st, &r)) L::interact(r2, dx, dy, dz, st, r, acc);
      }
    } else if (L::test(r2, st, &r)) {
      L::interact(r2, dx, dy, dz, st, r, acc);
    }
  };
  while (any()) {
    int q;
    const int j = locate(pop(), q);
    eval(j, q, rec[j]);
  }
}

// One i-chunk (TILE: one tile of the work queue) against a list of
// neighbour cells.  `stencil` selects the 27 neighbours of the i-chunk's
// cell; otherwise the single cell `cj0` with translation (tx0, ty0, tz0).
// i-chunks come from the chunk table itself (CI == CJ) unless TILE.
template <class L, int CI, int CJ, bool SUBSET, bool BUILD, bool TILE = false>
__device__ __forceinline__ void run_chunk(
    const DevParts &d, const LoopArgs &la, const ChunkArgs &ca,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    int ich, bool stencil, int cj0, double tx0, double ty0, double tz0,
    const NbrList &nl, const TileArgs *tq = nullptr, int tile = 0,
    const BundleNbr *nbl = nullptr, int nnb = 0) {
  using Sh = GatherLds<CI, CJ>;
  constexpr int S = Sh::S, RPC = Sh::RPC;
  constexpr int C = CI;
  constexpr int G = 64 / CJ;  // chunks per staging group
  static_assert(TILE || CI == CJ, "i-chunks are chunks of the table");
  static_assert(!BUILD || (CI == kListChunk && CJ == kListChunk &&
                           !L::kSymmetric),
                "lists are built by the density loop in list tiles");
  __shared__ Sh sh;
  const int lane = threadIdx.x & 63;
  const int g = lane % C;
  const int s = lane / C;
  // i: the lane's particle (field slot); ic: its record slot, the index
  // space of the chunk tables (the two differ with the time-bin order)
  int i, ic, cell_i;
  bool act;
  if (TILE) {
    // the tile builder applied the loop's predicate (and filt narrows it)
    const int tl = tq->ids ? uniform(tq->ids[tile]) : tile;
    const int ts = uniform(tq->tile_start[tl]);
    act = g < uniform(tq->tile_start[tl + 1]) - ts;
    i = act ? tq->alist[ts + g] : 0;
    ic = act ? rec_slot(d, i) : 0;
    if (tq->filt && act) act = tq->filt[i] != 0;
    cell_i = uniform(tq->tile_cell[tl]);
  } else {
    ic = ca.chunk_start[ich] + g;
    act = ic < ca.chunk_start[ich + 1];
    i = act ? field_slot(d, ic) : 0;
    cell_i = ca.chunk_cell[ich];
    if (act) {
      act = is_active(d.tb[i], la.max_active_bin);
      if (SUBSET) act = act && la.flag[i];
      // rebuild only the lists that do not cover the current radius
      if (BUILD && SUBSET) act = act && !(search_radius2(d.h[i]) <= d.r2b[i]);
    }
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
  if (act && s == 0 && la.fresh)
    L::fresh_acc(st, acc);
  else if (act && s == 0)
    L::load_acc(d, i, acc);
  else
    L::identity(acc);
  // squared radius of the candidate test (skinned when building lists)
  const float rt2 = BUILD ? nl.lam2 * st.hg2 : st.hg2;
  // pre-test radius; inactive lanes never pass
  const float rt2m = act ? pretest_radius2(rt2, la.pos_eps) : -1.f;
  const float hg2i = wave_max(act ? rt2 : 0.f);
  int ncnt = 0;  // list entries found by this lane (BUILD)

  // Neighbour table: lane q < nq describes neighbour q (the 27-cell
  // stencil, one cell, or the nnb <= 27 images of a task bundle item).
  const int nq = nbl ? nnb : (stencil ? 27 : 1);
  int nchq = 0;
  bool qself = false;
  if (lane < nq) {
    double tx = tx0, ty = ty0, tz = tz0;
    int cj = cj0;
    int cx, cy, cz;
    cell_coords(cell_i, la.grid, cx, cy, cz);
    if (nbl) {
      cj = nbl[lane].cell;
      tx = nbl[lane].t[0];
      ty = nbl[lane].t[1];
      tz = nbl[lane].t[2];
    } else if (stencil) {
      cj = neighbour(la.grid, cx, cy, cz, kStencilOrder[lane], tx, ty, tz);
    }
    int jx, jy, jz;
    cell_coords(cj, la.grid, jx, jy, jz);
    const double w[3] = {la.grid.dim[0] / la.grid.nc[0],
                         la.grid.dim[1] / la.grid.nc[1],
                         la.grid.dim[2] / la.grid.nc[2]};
    const double t[3] = {tx, ty, tz};
    const int cj3[3] = {jx, jy, jz};
    for (int k = 0; k < 3; k++) {
      sh.qorg[3 * lane + k] = (double)cj3[k] * w[k] + t[k];
    }
    const int k0 = ca.cell_chunk[cj], k1 = ca.cell_chunk[cj + 1];
    sh.qk0[lane] = k0;
    sh.qj0[lane] = la.cell_start[cj];
    sh.qj1[lane] = la.cell_start[cj + 1];
    sh.qcell[lane] = cj;
    sh.ttab[3 * lane] = tx;
    sh.ttab[3 * lane + 1] = ty;
    sh.ttab[3 * lane + 2] = tz;
    sh.ocode[lane] = stencil ? (unsigned int)kStencilOrder[lane] : 13u;
    nchq = k1 - k0;
  }
  int incl = nchq;  // inclusive prefix over lanes 0..31
#pragma unroll
  for (int off = 1; off < 32; off <<= 1) {
    const int v = __shfl_up(incl, off);
    if (lane >= off) incl += v;
  }
  if (lane < nq) sh.qpre[lane] = incl - nchq;
  const int total = __shfl(incl, nq - 1);
  // i-chunk box (a tile's box comes from its lanes)
  double ilx, ily, ilz, ihx, ihy, ihz;
  if (TILE) {
    ilx = act ? st.x : 1e300;
    ily = act ? st.y : 1e300;
    ilz = act ? st.z : 1e300;
    ihx = act ? st.x : -1e300;
    ihy = act ? st.y : -1e300;
    ihz = act ? st.z : -1e300;
#pragma unroll
    for (int off = 1; off < C; off <<= 1) {
      ilx = fmin(ilx, __shfl_xor(ilx, off));
      ily = fmin(ily, __shfl_xor(ily, off));
      ilz = fmin(ilz, __shfl_xor(ilz, off));
      ihx = fmax(ihx, __shfl_xor(ihx, off));
      ihy = fmax(ihy, __shfl_xor(ihy, off));
      ihz = fmax(ihz, __shfl_xor(ihz, off));
    }
    // every lane holds the tile's box: keep it in scalar registers
    ilx = uniform(ilx);
    ily = uniform(ily);
    ilz = uniform(ilz);
    ihx = uniform(ihx);
    ihy = uniform(ihy);
    ihz = uniform(ihz);
  } else {
    const double *bb = ca.bbox + 6 * ich;
    ilx = bb[0];
    ily = bb[1];
    ilz = bb[2];
    ihx = bb[3];
    ihy = bb[4];
    ihz = bb[5];
  }
  // FP16 pre-test frame: the centre of the i-box, lengths in units of the
  // largest search radius R of the tile.  Every rounding between the
  // positions and the half-precision squared distance is bounded by a few
  // units of 2^-11 of the coordinates' magnitude (at most V + 1.1 for the
  // candidates within reach, V the largest |i coordinate|), so widening
  // each radius r to (r + 2e)^2 (1 + 8 2^-11), e = 2 (V + 1.1) 2^-11, never
  // rejects a neighbour of the pair contract.
  // (the force loop's symmetric test needs a widened radius per candidate
  // as well, which costs more than the packed rows save: it stays FP32)
  constexpr bool kF16 = (RPC % 2 == 0) && !L::kSymmetric;
  constexpr float kU16 = 1.f / 2048.f;
  float invR = 0.f, eps16 = 0.f;
  _Float16 ixh = 0, iyh = 0, izh = 0, rt2h = -1;
  if (kF16) {
    const double cx = 0.5 * (ilx + ihx), cy = 0.5 * (ily + ihy),
                 cz = 0.5 * (ilz + ihz);
    invR = 1.f / sqrtf(hg2i);
    if (lane < nq) {
      sh.qoffc[3 * lane] = (float)(sh.qorg[3 * lane] - cx);
      sh.qoffc[3 * lane + 1] = (float)(sh.qorg[3 * lane + 1] - cy);
      sh.qoffc[3 * lane + 2] = (float)(sh.qorg[3 * lane + 2] - cz);
    }
    const float fx = (float)(st.x - cx) * invR;
    const float fy = (float)(st.y - cy) * invR;
    const float fz = (float)(st.z - cz) * invR;
    ixh = (_Float16)fx;
    iyh = (_Float16)fy;
    izh = (_Float16)fz;
    const float vbound =
        1.001f * wave_max(act ? fmaxf(fabsf(fx), fmaxf(fabsf(fy), fabsf(fz)))
                              : 0.f);
    eps16 = 2.f * (vbound + 1.1f) * kU16 + 1e-6f;
    const float r = sqrtf(rt2 * invR * invR) + 2.f * eps16;
    const float b = r * r * (1.f + 8.f * kU16) * (1.f + 1.f / 1024.f);
    rt2h = act ? (_Float16)fminf(b, 65504.f) : (_Float16)(-1.f);
  }
  wave_sync();
  // pair-contract frames of the neighbours (from the table in LDS)
  if (lane < nq) {
    int cx, cy, cz, jx, jy, jz;
    cell_coords(cell_i, la.grid, cx, cy, cz);
    cell_coords(sh.qcell[lane], la.grid, jx, jy, jz);
    const int ci3[3] = {cx, cy, cz}, cj3[3] = {jx, jy, jz};
    const double w[3] = {la.grid.dim[0] / la.grid.nc[0],
                         la.grid.dim[1] / la.grid.nc[1],
                         la.grid.dim[2] / la.grid.nc[2]};
    const double t[3] = {sh.ttab[3 * lane], sh.ttab[3 * lane + 1],
                         sh.ttab[3 * lane + 2]};
    int o[3];
    for (int k = 0; k < 3; k++) {
      int v = cj3[k] - ci3[k];
      v = v > 1 ? v - la.grid.nc[k] : (v < -1 ? v + la.grid.nc[k] : v);
      o[k] = v;
    }
    double fo[6];
    qself = pair_frame(o, ci3, cj3, t, w, fo);
    for (int k = 0; k < 6; k++) sh.fo[6 * lane + k] = fo[k];
  }
  const unsigned int selfm =
      (unsigned int)uniform((int)(unsigned int)__ballot(qself));
  wave_sync();

  unsigned long long m0 = 0ull, m1 = 0ull;  // row hits of this lane
  int nrow = 0;                 // rows recorded (uniform, multiple of RPC)
  for (int fb = 0; fb < total; fb += 64) {
    // Cull: one candidate chunk per lane.
    const int f = fb + lane;
    bool pass = false;
    unsigned int pe = 0u;
    if (f < total) {
      int q = 0;
#pragma unroll
      for (int b = 16; b > 0; b >>= 1)
        if (q + b < nq && sh.qpre[q + b] <= f) q += b;
      const int kc = sh.qk0[q] + (f - sh.qpre[q]);
      const double tx = sh.ttab[3 * q], ty = sh.ttab[3 * q + 1],
                   tz = sh.ttab[3 * q + 2];
      const double *jb = ca.bbox + 6 * kc;
      const double gx = fmax(0.0, fmax(jb[0] + tx - ihx, ilx - (jb[3] + tx)));
      const double gy = fmax(0.0, fmax(jb[1] + ty - ihy, ily - (jb[4] + ty)));
      const double gz = fmax(0.0, fmax(jb[2] + tz - ihz, ilz - (jb[5] + tz)));
      const float gap2 = (float)(gx * gx + gy * gy + gz * gz);
      const float rad2 = L::kSymmetric ? fmaxf(hg2i, ca.hg2max[kc]) : hg2i;
      pass = gap2 <= rad2 * 1.0001f;
      pe = ((unsigned int)q << 27) | (unsigned int)(kc - sh.qk0[q]);
    }
    const unsigned long long pm = __ballot(pass);
    if (pm == 0ull) continue;  // uniform
    const int np = __popcll(pm);
    if (pass) sh.plist[__popcll(pm & ((1ull << lane) - 1ull))] = pe;
    wave_sync();
    for (int pb = 0; pb < np; pb += G) {
      // Stage the candidates of up to G surviving chunks.
      {
        const int c = lane / CJ;
        if (pb + c < np) {
          const unsigned int e = sh.plist[pb + c];
          const int q = (int)(e >> 27);
          const int j = sh.qj0[q] + (int)(e & kRowMask) * CJ + lane % CJ;
          PosRec p;
          if (j < sh.qj1[q]) {
            p = pos[j];
          } else {  // past the end of a short chunk: never passes
            p.x = p.y = p.z = -1e30f;
            p.hg2m = -1.f;
          }
          if (kF16) {
            // slot of row pair (rr / 2) of stream sp, lane rr % 2 of the pair
            const int jl = lane % CJ, rr = jl / S, sp = jl % S;
            const int dst = ((c * (RPC / 2) + (rr >> 1)) * S + sp) * 2 + (rr & 1);
            sh.hx[dst] = (_Float16)((p.x + sh.qoffc[3 * q]) * invR);
            sh.hy[dst] = (_Float16)((p.y + sh.qoffc[3 * q + 1]) * invR);
            sh.hz[dst] = (_Float16)((p.z + sh.qoffc[3 * q + 2]) * invR);
          } else {
            sh.lpos[lane] = p;
          }
        }
      }
      wave_sync();
      for (int c = 0; c < G; c++) {
        if (pb + c >= np) break;  // uniform
        const unsigned int e = (unsigned int)uniform((int)sh.plist[pb + c]);
        const int q = (int)(e >> 27);
        const int js = uniform(sh.qj0[q]) + (int)(e & kRowMask) * CJ;
        const int nj = min(CJ, uniform(sh.qj1[q]) - js);
        // chunk slot of the lane's own particle (a particle is not its own
        // neighbour); out of the chunk's range unless the chunk holds it
        const int selfl = uniform(sh.qcell[q]) == cell_i ? ic - js : -1;
        // i relative to the neighbour image's corner
        const float xs = (float)(st.x - sh.qorg[3 * q]);
        const float ys = (float)(st.y - sh.qorg[3 * q + 1]);
        const float zs = (float)(st.z - sh.qorg[3 * q + 2]);
        if (lane == 0)
          sh.cinfo[nrow / RPC] = ((unsigned int)q << 27) | (unsigned int)js;
        // hits of this chunk's rows, bit rr for row rr
        using Nib = typename std::conditional<(RPC > 32), unsigned long long,
                                              unsigned int>::type;
        Nib nib = 0;
        if (kF16) {
          const half2v ix2 = {ixh, ixh}, iy2 = {iyh, iyh}, iz2 = {izh, izh};
#pragma unroll
          for (int pr = 0; pr < RPC / 2; pr++) {
            if (2 * pr * S >= nj) break;  // uniform
            const int idx = 2 * ((c * (RPC / 2) + pr) * S + s);
            const half2v x2 = *reinterpret_cast<const half2v *>(&sh.hx[idx]);
            const half2v y2 = *reinterpret_cast<const half2v *>(&sh.hy[idx]);
            const half2v z2 = *reinterpret_cast<const half2v *>(&sh.hz[idx]);
            const half2v dx = ix2 - x2, dy = iy2 - y2, dz = iz2 - z2;
            const half2v r2 = dz * dz + (dy * dy + dx * dx);
            const bool p0 = r2.x < rt2h, p1 = r2.y < rt2h;
            const int jl0 = 2 * pr * S + s, jl1 = jl0 + S;
            nib |= (p0 && jl0 != selfl) ? ((Nib)1 << (2 * pr)) : (Nib)0;
            nib |= (p1 && jl1 != selfl) ? ((Nib)1 << (2 * pr + 1)) : (Nib)0;
          }
        } else {
#pragma unroll
          for (int rr = 0; rr < RPC; rr++) {
            if (rr * S >= nj) break;  // uniform
            const int jl = rr * S + s;
            const PosRec p = sh.lpos[c * CJ + jl];
            const float dx = xs - p.x, dy = ys - p.y, dz = zs - p.z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            const bool pass =
                L::kSymmetric ? (r2 < rt2m || r2 < p.hg2m) : r2 < rt2m;
            nib |= (pass && jl != selfl) ? ((Nib)1 << rr) : (Nib)0;
          }
        }
        if (nrow < 64)  // uniform
          m0 |= (unsigned long long)nib << nrow;
        else
          m1 |= (unsigned long long)nib << (nrow - 64);
        nrow += RPC;
        if (nrow == 64 * kMasks) {
          wave_sync();
          interact_rows<L, CI, CJ, BUILD>(rec, sh, st, s, m0, m1, acc, nl, i,
                                          rt2, ncnt, selfm);
          m0 = m1 = 0ull;
          nrow = 0;
        }
      }
      wave_sync();
    }
  }
  if (nrow > 0) {
    wave_sync();
    interact_rows<L, CI, CJ, BUILD>(rec, sh, st, s, m0, m1, acc, nl, i, rt2,
                                    ncnt, selfm);
  }
  L::template combine<S>(acc);
  if (act && s == 0) L::store_acc(d, i, acc);
  if (BUILD) {
    int most = ncnt;
#pragma unroll
    for (int off = 64 / S; off < 64; off <<= 1)
      most = max(most, __shfl_xor(most, off));
    if (act) nl.cnt[i * kListStreams + s] = ncnt;
    // a list with an overflowing stream is unusable
    if (act && s == 0) d.r2b[i] = most > nl.Ks ? -1.f : rt2;
  }
}

// Batched: every chunk against its 27-cell stencil.
template <class L, int C, bool SUBSET, bool BUILD>
__global__ SWSPH_LOOP_BOUNDS(L) void k_loop_stencil(
    DevParts d, LoopArgs la, ChunkArgs ca,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    NbrList nl) {
  const int ich = xcd_item(blockIdx.x, threadIdx.x >> 6, ca.nchunks);
  if (ich >= ca.nchunks) return;  // uniform per wave
  run_chunk<L, C, C, SUBSET, BUILD>(d, la, ca, rec, pos, ich, true, 0, 0.0,
                                    0.0, 0.0, nl);
}

// Task bundle: one wave per item (an i-chunk and its neighbour images, or
// with n < 0 its cell's whole 27-cell stencil).
template <class L, int C>
__global__ __launch_bounds__(64) void k_loop_bundle(
    DevParts d, LoopArgs la, ChunkArgs ca,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    BundleArgs ba) {
  const BundleItem it = ba.items[blockIdx.x];
  const NbrList none{};
  const int nn = uniform(it.n);
  run_chunk<L, C, C, false, false, false>(
      d, la, ca, rec, pos, uniform(it.ich), nn < 0, 0, 0.0, 0.0, 0.0, none,
      nullptr, 0, nn < 0 ? nullptr : ba.nbrs + uniform(it.off),
      nn < 0 ? 0 : nn);
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
  const NbrList none{};
  run_chunk<L, C, C, SUBSET, false, false>(d, la, ca, rec, pos, ich, false,
                                              ta.jcell[q], ta.t[3 * q],
                                              ta.t[3 * q + 1], ta.t[3 * q + 2],
                                              none);
}

// Persistent form: a fixed grid of waves drains the tile queue.  A wave
// starts on the queue of its XCD (blocks are dealt round-robin over the 8
// XCDs), takes kGrab tiles per atomic, and moves on to the next queue when
// its own is exhausted; every queue is visited at most once after it runs
// dry, so every wave exits.
constexpr int kGrab = 2;
template <class L, int CI, int CJ, bool BUILD>
__global__ __launch_bounds__(64, kOccSparseQueue) void k_loop_queue(
    DevParts d, LoopArgs la, ChunkArgs ca,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    NbrList nl, TileArgs tq) {
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
      run_chunk<L, CI, CJ, false, BUILD, true>(d, la, ca, rec, pos, 0, true, 0,
                                               0.0, 0.0, 0.0, nl, &tq, t);
    }
  }
}

// Grid form of the tile queue: one wave per tile over a grid sized for the
// largest possible tile count; waves past the device-side count exit.
template <class L, int CI, int CJ, bool BUILD>
__global__ SWSPH_TILE_BOUNDS(L) void k_loop_tiles(
    DevParts d, LoopArgs la, ChunkArgs ca,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    NbrList nl, TileArgs tq) {
  const int t = xcd_item(blockIdx.x, threadIdx.x >> 6, *tq.ntiles);
  if (t >= *tq.ntiles) return;  // uniform per wave
  run_chunk<L, CI, CJ, false, BUILD, true>(d, la, ca, rec, pos, 0, true, 0, 0.0,
                                           0.0, 0.0, nl, &tq, t);
}

// The tiles that a grid of b0 blocks of k_loop_tiles does not reach (a
// grid sized for an earlier tile count, as in a captured graph): blocks
// b0, b0 + 1, ... of the same block -> tile mapping, visited in strides of
// this kernel's grid.  Normally there are none and every wave exits at
// once; the loop around the tile costs registers, so the main grid runs
// one tile per wave.
template <class L, int CI, int CJ, bool BUILD>
__global__ __launch_bounds__(64) void k_loop_tiles_rest(
    DevParts d, LoopArgs la, ChunkArgs ca,
    const typename L::Rec *__restrict__ rec, const PosRec *__restrict__ pos,
    NbrList nl, TileArgs tq, int b0) {
  const int n = *tq.ntiles;
  const int nb = xcd_blocks(n);
  for (int b = b0 + (int)blockIdx.x; b < nb; b += (int)gridDim.x) {
    const int t = xcd_item(b, threadIdx.x >> 6, n);
    if (t >= n) continue;  // uniform per wave
    wave_sync();
    run_chunk<L, CI, CJ, false, BUILD, true>(d, la, ca, rec, pos, 0, true, 0,
                                             0.0, 0.0, 0.0, nl, &tq, t);
  }
}

// List loop: lane g + kListChunk * s walks stream s of the neighbour list of
// particle g of its tile (COMPACT: of d.sel[wave * kListChunk + g]),
// re-evaluates the pair contract for each entry and interacts with the ones
// that pass the loop's own test.  Particles whose list does not cover their
// search radius are left to the chunk loop.
template <class L, bool COMPACT>
__global__ SWSPH_LOOP_BOUNDS(L) void k_loop_list(
    DevParts d, LoopArgs la, ChunkArgs ca, NbrList nl,
    const typename L::Rec *__restrict__ rec, int nsel) {
  constexpr int C = kListChunk, S = kListStreams;
  static_assert(!L::kSymmetric, "lists hold gather neighbours only");
  const int lane = threadIdx.x & 63;
  const int g = lane % C;
  const int s = lane / C;
  int i = 0;
  bool act;
  if (COMPACT) {
    const int k = (blockIdx.x * kWaves + (threadIdx.x >> 6)) * C + g;
    act = k < nsel;
    if (act) i = d.sel[k];
  } else {
    const int ich = xcd_item(blockIdx.x, threadIdx.x >> 6, ca.nchunks);
    if (ich >= ca.nchunks) return;
    const int j = ca.chunk_start[ich] + g;  // record slot
    act = j < ca.chunk_start[ich + 1];
    if (act) i = field_slot(d, j);
  }
  if (act) act = is_active(d.tb[i], la.max_active_bin);
  typename L::I st;
  if (act) {
    L::load_i(d, i, st);
    act = st.hg2 <= d.r2b[i];
  }
  if (!__any(act)) return;  // uniform
  typename L::Acc acc;
  if (act && s == 0 && la.fresh)
    L::fresh_acc(st, acc);
  else if (act && s == 0)
    L::load_acc(d, i, acc);
  else
    L::identity(acc);
  const int n = act ? min(nl.cnt[i * S + s], nl.Ks) : 0;
  const unsigned int *ip = nl.idx + (size_t)i * S + s;
  const size_t stride = (size_t)nl.npad * S;
  // the i-cell's coordinates
  int ci3[3] = {0, 0, 0};
  if (act) cell_coords(d.cell[i], la.grid, ci3[0], ci3[1], ci3[2]);
  const double w[3] = {la.grid.dim[0] / la.grid.nc[0],
                       la.grid.dim[1] / la.grid.nc[1],
                       la.grid.dim[2] / la.grid.nc[2]};
  // Software pipeline: list entries are fetched two ahead and records one
  // ahead of the entry being evaluated, so the gathers overlap the bodies.
  unsigned int e1 = n > 0 ? ip[0] : 0u;
  unsigned int e2 = n > 1 ? ip[stride] : 0u;
  typename L::Rec r1;
  if (n > 0) r1 = rec[e1 & kRowMask];
  for (int k = 0; k < n; k++) {
    const unsigned int e = e1;
    const typename L::Rec jr = r1;
    e1 = e2;
    if (k + 2 < n) e2 = ip[(size_t)(k + 2) * stride];
    if (k + 1 < n) r1 = rec[e1 & kRowMask];
    // stencil offset code of j's cell: the neighbour cell, its image
    // translation and the pair-contract frame
    const int oc = (int)(e >> 27);
    double t[3];
    const int cj = neighbour(la.grid, ci3[0], ci3[1], ci3[2], oc, t[0], t[1],
                             t[2]);
    int cj3[3];
    cell_coords(cj, la.grid, cj3[0], cj3[1], cj3[2]);
    const int o[3] = {oc / 9 - 1, (oc / 3) % 3 - 1, oc % 3 - 1};
    double fo[6];
    const bool self = pair_frame(o, ci3, cj3, t, w, fo);
    float dx, dy, dz;
    const float r2 =
        pair_r2(self, st.x, st.y, st.z, jr.x, jr.y, jr.z, fo, dx, dy, dz);
    if (L::test(r2, st, &jr)) L::interact(r2, dx, dy, dz, st, jr, acc);
  }
  L::template combine<S>(acc);
  if (act && s == 0) L::store_acc(d, i, acc);
}

}  // namespace sphl
