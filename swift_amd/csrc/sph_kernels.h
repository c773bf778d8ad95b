// sph_kernels.h -- device data layout and launchers shared by the engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Neighbour records: the fields a neighbour loop reads from particle j,
// packed so that the interaction phase fetches them with a few wide
// loads.  The leading double position (and, for the force loop, the search
// radius) is what the candidate test reads.  Inhibited particles get a far
// position (and a negative radius) so that every test rejects them.
// Record sizes: the hit walks gather records of scattered j from L2, and
// fewer bytes per record measured faster than cache-line-aligned strides
// (64 B density / 128 B force records: +6 % / +5 % loop time).  Records are
// 8-byte aligned with no padding: density 40 B, gradient 56 B, and a force
// record without h^-4 (two multiplies per pair instead), 96 B (-2.7 % per
// pass against 48 / 64 / 112 B records with h^-4 and padding).
struct __attribute__((aligned(8))) DensRec {
  double x, y, z;
  float vx, vy, vz, m;
};
struct __attribute__((aligned(8))) GradRec {
  double x, y, z;
  float vx, vy, vz, mrho;  // m_j / rho_j
  float u, c, alpha, pad0;
};
struct __attribute__((aligned(16))) ForceRec {
  double x, y, z;
  float hg2;               // (gamma_K h_j)^2
  float vx, vy, vz;
  float m, inv_m, h_inv;
  float rho, inv_rho, pr2, f;  // pr2 = P / rho^2
  float bal, c, alpha, P;
  float Pad, u, tb;            // Pad = P * alpha_diff
};

// Candidate record of the FP32 pre-test: the position relative to the
// corner of the particle's top-level cell, rounded to float, and the
// particle's search radius (gamma_K h)^2 widened by the pre-test margin
// (read by the force test).  Inhibited particles are far away with a
// negative radius.
struct __attribute__((aligned(16))) PosRec {
  float x, y, z, hg2m;
};

// Chunks: runs of <= C consecutive particles of one cell (Morton order inside
// the cell), with their bounding box and largest search radius.
struct ChunkArgs {
  const int *chunk_start;   // [nchunks + 1] first particle of each chunk
  const int *cell_chunk;    // [ncells + 1] first chunk of each cell
  const int *chunk_cell;    // [nchunks] cell of each chunk
  const double *bbox;       // [6 nchunks] lo x,y,z, hi x,y,z
  const float *hg2max;      // [nchunks] max (gamma_K h)^2
  // [2 nchunks] the box of the chunk's pre-test records (PosRec x, y, z
  // relative to the cell corner) with their largest widened radius in the
  // first .w, or null (tables the sparse tiles do not read)
  const float4 *fbox;
  int nchunks;
  int C;                    // chunk capacity (8, 16, 32 or 64)
};

// Work queue of the persistent loops: tiles of <= C particles of one cell,
// cut from a compact, cell-ordered list of the particles a loop processes
// (all active ones, or the flagged ones), built on the device.  Tile t
// holds alist[tile_start[t] .. tile_start[t + 1]).  Waves take tiles from
// eight queues (one per XCD, each over a contiguous range of tiles) through
// the counters heads[64 q], and steal from the other queues when their own
// runs dry.
struct TileArgs {
  const int *alist;
  const int *tile_start;  // [ntiles + 1]
  const int *tile_cell;   // [ntiles]
  const int *ntiles;      // device-side count (of ids, when ids is set)
  int *heads;             // [8 * 64]
  // A subset of the tiles: work item t is tile ids[t], and only the
  // particles with filt[i] != 0 take part (null: every tile, particle).
  const int *ids;
  const int *filt;
};

// Structure-of-arrays particle store in HBM, sorted by top-level cell.
struct DevParts {
  double *px, *py, *pz;
  float *vx, *vy, *vz, *m, *u, *h;
  int8_t *tb;
  // density sums (raw until the ghost finalises them)
  float *rho, *rho_dh, *wc, *wc_dh, *divv, *rotx, *roty, *rotz;
  int *nd;
  // ghost state
  float *left, *right;
  int *redo;
  // prepared by the ghost (hydro_prepare_gradient / hydro_reset_gradient)
  float *P, *c, *f, *bal;
  // gradient sums
  float *vsig, *lapu, *avmax;
  int *ng;
  // viscosity / diffusion state (hydro_prepare_force)
  float *alpha, *adiff, *divv_prev, *divv_dt;
  // force sums
  float *ax, *ay, *az, *udt, *hdt;
  int *nf;
  int8_t *mintb;
  // neighbour records (rebuilt after the stage that changes their fields)
  DensRec *rd;
  GradRec *rg;
  ForceRec *rf;
  PosRec *pos;
  // neighbour lists: squared radius each particle's list was built with
  // (negative: no usable list), gradient fall-back flags, and the compact
  // list of particles the ghost sent back for another density round
  float *r2b;
  int *lfb;
  int *sel;
  int *cell;  // top-level cell of each particle
  // time integration (struct xpart: v_full, u_full) and the caller's index
  // of every slot
  float *vfx, *vfy, *vfz, *ufull;
  int *perm;
  // With the time-bin order the fields are in (time bin, Morton) order
  // inside each cell while the records (rd, rg, rf, pos) and the chunk
  // tables keep Morton order (the loops' candidates stay spatially
  // compact): cand[i] is the record slot of field slot i, soa[j] the field
  // slot of record slot j.  Both null: one order.  Cells cover the same
  // slot ranges in both.
  int *cand, *soa;
};

// Record slot of field slot i, and field slot of record slot j.
__device__ __forceinline__ int rec_slot(const DevParts &d, int i) {
  return d.cand ? d.cand[i] : i;
}
__device__ __forceinline__ int field_slot(const DevParts &d, int j) {
  return d.soa ? d.soa[j] : j;
}

// Per-particle neighbour lists (Verlet lists with a skin): the candidates
// j with r_ij^2 < lam2 * (gamma_K h_i)^2 found by the chunk loop, kept for
// the density re-runs of the ghost and the gradient loop, which then test
// only these.  Stream s (of kListStreams) of particle i owns slots
// idx[(k * npad + i) * kListStreams + s], k < cnt[i * kListStreams + s];
// an entry is (offset code << 27) | j with offset code
// 9 (ox + 1) + 3 (oy + 1) + (oz + 1) for the stencil offset (ox, oy, oz) of
// j's cell from i's cell.
constexpr int kListStreams = 4;
constexpr int kListChunk = 64 / kListStreams;
// Smallest tile (i-particles per wave) of the work-queue schedules; the
// tile tables are sized for it.
constexpr int kMinTile = 4;
// Slots of the ghost's density-work counter (launch_ghost): the counter
// block holds 3 + kGhostWorkSlots 64-bit counters.
constexpr int kGhostWorkSlots = 32;
constexpr int kGhostCounters = 3 + kGhostWorkSlots;
// Sparse tiles (sph_sparse.h): tiles of kMinTile particles, candidates
// culled by the chunks of kSparseChunk, then by their sub-chunks of
// kSparseSub against each particle.
constexpr int kSparseChunk = 32;
constexpr int kSparseSub = 8;
// Largest particle and cell counts of the allocation-free batched pass,
// whose per-pass scratch is a fixed device array (sph_scratch.hip).
constexpr int64_t kNoAllocMaxParts = (int64_t)1 << 21;
constexpr int kNoAllocMaxCells = 1 << 16;
struct NbrList {
  unsigned int *idx;
  int *cnt;
  int Ks;      // slots per stream
  int npad;
  float lam2;  // skin factor on the squared search radius
};

struct GridDesc {
  int nc[3];
  double dim[3];
};

struct LoopArgs {
  GridDesc grid;
  const int *cell_start;  // [ncells + 1]
  const int *flag;        // subset loops: particles with flag[i] != 0
  int max_active_bin;
  // Bound on the per-component difference between the FP32 pre-test
  // distance and the pair contract's distance (see pretest_radius2).
  float pos_eps;
  // Work-queue selections take only particles of cells with cmask[c] != 0
  // (null: every cell) -- the cells of a bundle of cell tasks.
  const unsigned char *cmask;
  // 1: every particle a launch takes meets its whole 27-cell stencil in one
  // wave, so its accumulators start from their initial values in registers
  // (L::fresh_acc) and are stored, not loaded (the batched loops; their
  // resets are skipped).
  int fresh;
};

// Up to two directed cell interactions (i-cell gathers from j-cell, the
// image of the j-cell being translated by t).
struct TaskArgs {
  int icell[2];
  int jcell[2];
  int ntiles[2];
  double t[6];
};

// A bundle of released self / pair tasks, merged per i-chunk: item k runs
// i-chunk items[k].ich against the neighbour images nbrs[off .. off + n)
// (at most 27: the j-cell and the translation applied to its positions).
struct BundleNbr {
  int cell, pad;
  double t[3];
};
// n < 0: the i-chunk's cell meets all 27 images of its stencil (no list).
struct BundleItem {
  int ich, off, n, pad;
};
struct BundleArgs {
  const BundleItem *items;
  const BundleNbr *nbrs;
  int nitems;
};

struct GhostArgs {
  float eta_dim, eps, h_max, h_min;
  int max_active_bin;
  double time_base;
  float alpha_max, alpha_min, visc_length;
  float diff_beta, diff_alpha_max, diff_alpha_min;
  // 1: the force loop that follows starts its sums from the identity, so
  // the extra ghost leaves the force accumulators alone.
  int fresh_force;
  // 1: the gradient loop that follows starts its sums from the particle's
  // own v_sig and alpha, so the ghost leaves the gradient accumulators alone.
  int fresh_grad;
  // 1: neighbour lists are in use (the ghost keeps their fallback flags).
  int lists;
  // 1: the ghost writes the gradient record of each particle it finishes,
  // the extra ghost the force record of each particle (no separate builds).
  int grad_rec, force_rec;
};

enum class LoopKind { Density, Gradient, Force };

// Drift of particles [lo, hi) from ti_old to ti_current (intervals in
// physical time; without cosmology all three are equal).
struct DriftArgs {
  double dt_drift, dt_kick_hydro, dt_therm;
  float min_u;    // hydro_props minimal_internal_energy
  GridDesc g;
  int lo, hi;
};

// Half-kicks of particles [lo, hi): mode 1 kick1, mode 2 kick2.
struct KickArgs {
  long long ti_current;
  double time_base;
  float min_u;
  int mode, max_active_bin, lo, hi;
};

// Time steps of the active particles of [lo, hi).
struct StepArgs {
  long long ti_current;
  double time_base_inv;
  double dt_min, dt_max;
  float cfl, log_max_h_change;
  int max_active_bin, fixed, lo, hi;
};

// Scratch of the tile builder, sized for n particles and ncells cells.
struct TileScratch {
  int *flag;        // [n]
  int *scan;        // [n + 1]
  int *blk;         // [n / 1024 + 2]
  int *cnt;         // [ncells]
  int *cscan;       // [ncells + 1]
  int *cblk;        // [ncells / 1024 + 2]
  int *alist;       // [n]
  int *tile_start;  // [n / C + ncells + 1] for the smallest C in use
  int *tile_cell;
  int *ntiles;      // [1]
  int *heads;       // [8 * 64]
};
namespace sphk {
hipError_t launch_stencil(LoopKind kind, bool subset, const DevParts &d,
                          const LoopArgs &la, const ChunkArgs &ca,
                          hipStream_t st);
// Density loop over the 27-cell stencil that also (re)builds the neighbour
// lists of the particles it processes (all active ones, or with `subset`
// the flagged ones whose list does not cover their search radius).  The
// chunk table must have capacity kListChunk.
hipError_t launch_build_lists(bool subset, const DevParts &d,
                              const LoopArgs &la, const ChunkArgs &ca,
                              const NbrList &nl, hipStream_t st);
// A density or gradient loop over the neighbour lists of the active
// particles whose list covers their search radius: all of them (chunk
// table of capacity kListChunk), or the nsel particles d.sel[0..nsel).
hipError_t launch_list_loop(LoopKind kind, bool compact, const DevParts &d,
                            const LoopArgs &la, const ChunkArgs &ca,
                            const NbrList &nl, int nsel, hipStream_t st);
hipError_t launch_fill_f32(float *p, float v, int n, hipStream_t st);

// Tiles of <= C particles over the active particles (mode 0) or the active
// particles with flag[i] != 0 (mode 1); all device side, no host sync.
hipError_t launch_build_tiles(const DevParts &d, const LoopArgs &la, int n,
                              int ncells, int C, int mode, const int *flag,
                              const TileScratch &ts, hipStream_t st);
// A neighbour loop over the tiles (of <= CI particles) against candidate
// chunks of CJ particles (the table `ca`): by a persistent grid of `waves`
// waves (0: full occupancy on every CU) when `persistent`, else by one wave
// per tile over a grid for `max_tiles` tiles, and with `rest` a small grid
// that takes the tiles beyond the first max_tiles (a tile count that grew
// past the grid a captured graph was sized for).  Shapes: CJ = CI, or
// CJ = CI / 2 for CI = 16, 32, and the narrow tiles.
hipError_t launch_queue(LoopKind kind, bool build_lists, const DevParts &d,
                        const LoopArgs &la, const ChunkArgs &ca,
                        const TileArgs &ta, const NbrList &nl, int CI, int CJ,
                        bool persistent, int waves, int max_tiles, bool rest,
                        hipStream_t st);
// The tiles of the table in `ts` that hold a particle with filt[i] != 0,
// appended to ids (in no fixed order: each tile's sums are its own) with
// their count in *nids.
hipError_t launch_select_tiles(const TileScratch &ts, const int *filt,
                               int *ids, int *nids, hipStream_t st);
// The sparse form of the same loop over tiles of 4 particles: candidates
// culled by the chunks of 32 of `c1` against the tile, then by the chunks
// of 8 of `c2` (the same cells) against each particle (sph_sparse.h).
hipError_t launch_sparse(LoopKind kind, const DevParts &d, const LoopArgs &la,
                         const ChunkArgs &c1, const ChunkArgs &c2,
                         const TileArgs &ta, bool persistent, int waves,
                         int max_tiles, bool rest, hipStream_t st);
hipError_t launch_task(LoopKind kind, bool subset, const DevParts &d,
                       const LoopArgs &la, const ChunkArgs &ca,
                       const TaskArgs &ta, hipStream_t st);
// One launch for a task bundle: one wave per item.
hipError_t launch_bundle(LoopKind kind, const DevParts &d, const LoopArgs &la,
                         const ChunkArgs &ca, const BundleArgs &ba,
                         hipStream_t st);
// Positions / search radii of the candidate test for particles [lo, hi).
hipError_t launch_build_pos(const DevParts &d, const GridDesc &g, float pos_eps,
                            int lo, int hi, hipStream_t st);
// *moved = 1 if perm[s] != s for some s < n, else 0.
hipError_t launch_perm_moved(const int *perm, int n, unsigned int *moved,
                             hipStream_t st);
// Fills the per-chunk starts ([nchunks + 1]) and cells ([nchunks]) of a
// chunk table of capacity C from the cells' slot ranges (cell_start) and
// their first chunks (cell_chunk, [ncells + 1]); every cell holds
// ceil(count / C) chunks.
hipError_t launch_chunk_fill(const int *cell_start, const int *cell_chunk,
                             int ncells, int C, int nchunks, int n, int *start,
                             int *chunk_cell, hipStream_t st);
hipError_t launch_zero_u32(unsigned int *p, int n, hipStream_t st);
// mask[c] = 1 for the n listed cells, 0 for the other ncells - n.
hipError_t launch_cell_mask(unsigned char *mask, int ncells, const int *cells,
                            int n, hipStream_t st);
// act[c] = 1 for the cells that hold an active particle, else 0.
hipError_t launch_cell_active(const DevParts &d, int n, int ncells,
                              int max_active_bin, unsigned char *act,
                              hipStream_t st);
// inv[perm[s]] = s over the n slots.
hipError_t launch_invert_perm(const int *perm, int *inv, int n,
                              hipStream_t st);
// Subset flags of a task on the slots [lo, hi) of its cell: cleared, then
// set for the caller indices first + ind[k] (k < count, on the device);
// *err counts those whose slot is not in [lo, hi).
hipError_t launch_mark_subset(int *flag, const int *inv, long long first,
                              const int *ind, int count, int lo, int hi,
                              unsigned int *err, hipStream_t st);
// dst[i] = src[i] for i in [lo, hi) with (i - lo) % nt == off, else 0.
hipError_t launch_share_flags(int *dst, const int *src, int lo, int hi,
                              int off, int nt, hipStream_t st);
hipError_t launch_zero_u64(unsigned long long *p, int n, hipStream_t st);
// Bounding boxes and largest search radius of chunks [c0, c1); the largest
// search radius over them is max-reduced into *hg2_all (float bits).  With
// fbox, also the boxes of the chunks' pre-test records (grid g, pos_eps as
// for launch_build_pos).
hipError_t launch_chunk_bounds(const DevParts &d, const int *chunk_start,
                               int c0, int c1, int C, double *bbox,
                               float *hg2max, unsigned int *hg2_all,
                               hipStream_t st, const GridDesc &g, float pos_eps,
                               float4 *fbox);
// light: only the ghost's flags and brackets (the density loop that follows
// starts its sums from the identity).
// The radii of launch_chunk_bounds alone (positions unchanged since the
// boxes were formed), and with `records` the radii of the candidates'
// pre-test records (the table must cover every particle).
hipError_t launch_radius_bounds(const DevParts &d, const int *chunk_start,
                                int c0, int c1, int C, float *hg2max,
                                unsigned int *hg2_all, float pos_eps,
                                float4 *fbox, bool records, hipStream_t st);
hipError_t launch_init_density(const DevParts &d, int lo, int hi,
                               int max_active_bin, float h_max, hipStream_t st,
                               const int *list = nullptr,
                               const int *cnt = nullptr, bool light = false);
// counters[0]: particles sent back for another density round (their
// indices are appended to d.sel); [1]: of those, the ones whose list no
// longer covers the new radius; [2]: converged particles whose list does
// not cover their final radius (flagged in d.lfb; with lists only);
// [3, 3 + kGhostWorkSlots): density interactions evaluated for the
// processed particles, spread over the slots (one atomic per block on one
// address serialised the launch) and accumulated over the rounds: their
// sum is the count (the pass zeroes them).  [0..2] are zeroed by each
// launch.  Over a
// range, only the slots i with (i - lo) % nt == off are processed (one of
// nt ghost tasks of a cell, runner_do_ghost's offset / ntasks).
hipError_t launch_ghost(const DevParts &d, const GhostArgs &ga, int lo, int hi,
                        unsigned long long *counters, hipStream_t st,
                        const int *list = nullptr, const int *cnt = nullptr,
                        int off = 0, int nt = 1);
// The per-particle kernels below take either the particle range [lo, hi)
// or (list != nullptr) the particles list[0 .. *cnt) of a compact list whose
// capacity is hi - lo.
hipError_t launch_extra_ghost(const DevParts &d, const GhostArgs &ga, int lo,
                              int hi, hipStream_t st,
                              const int *list = nullptr,
                              const int *cnt = nullptr);
hipError_t launch_end_force(const DevParts &d, int lo, int hi,
                            int max_active_bin, hipStream_t st,
                            const int *list = nullptr,
                            const int *cnt = nullptr);
hipError_t launch_count(const DevParts &d, int n, int max_active_bin,
                        unsigned long long *out3, hipStream_t st);
// Copies n floats from each src[k] to dst[k], k < m <= kMaxCopies, in one
// launch (the pass-state save / restore).
constexpr int kMaxCopies = 16;
struct CopySet {
  const float *src[kMaxCopies];
  float *dst[kMaxCopies];
  int m;
};
hipError_t launch_copy_set(const CopySet &cs, int n, hipStream_t st);
// *out = number of active particles of [0, n) (device-side).
hipError_t launch_count_active(const DevParts &d, int n, int max_active_bin,
                               unsigned int *out, hipStream_t st);
// Records of particles [lo, hi); with max_active_bin >= 0 only those of the
// active particles.
hipError_t launch_build_records(LoopKind kind, const DevParts &d, int lo,
                                int hi, hipStream_t st, int max_active_bin = -1,
                                const int *list = nullptr,
                                const int *cnt = nullptr);

// ---- binning and time integration (sph_step.hip) ----
// Keys (cell << 30 | in-cell Morton code) of n positions, component k of
// particle i at p_k[stride * i]; val[i] = i.  With tb (the particles' time
// bins, same index), (cell << 30 + kBinBits) | (bin << 30) | Morton.
constexpr int kBinBits = 6;
hipError_t launch_bin_keys(const double *px, const double *py, const double *pz,
                           int stride, const GridDesc &g, const double inv_w[3],
                           int n, unsigned long long *key, int *val,
                           hipStream_t st, const int8_t *tb = nullptr);
// Stable radix sort of (key, value) pairs on the low `bits` key bits; with
// temp == nullptr only sets `bytes` (rocprim::radix_sort_pairs).
hipError_t sort_pairs(void *temp, size_t &bytes, const unsigned long long *kin,
                      unsigned long long *kout, const int *vin, int *vout,
                      int n, int bits, hipStream_t st);
hipError_t launch_wrap(const DevParts &d, const GridDesc &g, int n,
                       hipStream_t st);
// cell[k] and cell_start[0..ncells] from sorted keys (bin_bits: 0, or
// kBinBits for keys with the time bin).
hipError_t launch_cell_bounds(const unsigned long long *key, int n, int ncells,
                              int *cell_start, int *cell, hipStream_t st,
                              int bin_bits = 0);
// dst[k] = src[stride * idx[k] + comp]
template <class T>
hipError_t launch_gather(T *dst, const T *src, int stride, int comp,
                         const int *idx, int n, hipStream_t st);
// dst[stride * idx[k] + comp] = src[k]
template <class T>
hipError_t launch_scatter(T *dst, const T *src, int stride, int comp,
                          const int *idx, int n, hipStream_t st);
hipError_t launch_fill_i8(int8_t *p, int8_t v, int n, hipStream_t st);
hipError_t launch_drift(const DevParts &d, const DriftArgs &a,
                        unsigned int *dmax, hipStream_t st);
hipError_t launch_kick(const DevParts &d, const KickArgs &a, hipStream_t st);
hipError_t launch_timestep(const DevParts &d, const StepArgs &a,
                           unsigned long long *ti_end_min, unsigned int *err,
                           hipStream_t st);
// The fixed scratch of the allocation-free pass (sph_scratch.hip): its size,
// the one-time lookup of its address on `device` (which must be current),
// and the work queue and counters of a pass over n particles on ncells
// cells carved from it (hipErrorInvalidValue beyond the limits above,
// hipErrorNotInitialized before noalloc_setup on that device).
size_t noalloc_scratch_bytes();
hipError_t noalloc_setup(int device);
hipError_t noalloc_scratch(int device, int64_t n, int ncells, TileScratch *ts,
                           size_t *tile_cap, unsigned long long **redo,
                           unsigned long long **cnt);
}  // namespace sphk
