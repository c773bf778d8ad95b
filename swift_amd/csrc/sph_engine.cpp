// sph_engine.cpp -- host side of libswsph.so: the C ABI declared in
// include/swsph.h.
//
// The engine keeps the gas particles resident in HBM as a structure of
// arrays sorted by top-level cell (binned on the device, sph_step.hip) and
// issues the kernels of sph_kernels.hip and sph_step.hip on one HIP stream.  The per-task entry points are the drop-in for SWIFT's
// runner task functions; swsph_pass_plain drives them in the reference's
// dependency order (src/engine_maketasks.c:2546-2573: density -> ghost ->
// gradient -> extra ghost -> force -> end force), swsph_pass_batched runs
// the same pass as a few launches over all cells, and swsph_engine_step
// advances the gas in time around it (drift, kicks, time steps).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/swsph.h"
#include "sph_kernels.h"
#include "sph_sched.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(call)                                                       \
  do {                                                                     \
    hipError_t err_ = (call);                                              \
    if (err_ != hipSuccess)                                                \
      return fail(2, "%s failed: %s (%s:%d)", #call,                        \
                  hipGetErrorString(err_), __FILE__, __LINE__);            \
  } while (0)

template <class T>
hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
}

}  // namespace

struct swsph_engine {
  swsph_props props{};
  GridDesc grid{};
  int ncells = 0;
  int64_t n = 0;
  std::vector<int> cell_start;    // host copy [ncells+1]
  std::vector<int4> tiles_h;
  std::vector<int> cell_tiles;    // tiles per cell
  DevParts d{};
  int *d_cell_start = nullptr;
  int4 *d_tiles = nullptr;
  int ntiles = 0;
  unsigned long long *d_redo = nullptr;  // ghost counters (launch_ghost)
  unsigned long long *h_redo = nullptr;  // pinned readback
  int64_t density_work = 0;  // density interactions over all rounds
  unsigned long long *d_cnt = nullptr;
  unsigned long long *h_cnt = nullptr;
  hipStream_t st = nullptr;
  std::mutex mtx;
  bool timing = false;
  double t_ms[3] = {0, 0, 0};
  int64_t t_launch[3] = {0, 0, 0};
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  size_t ev_used = 0;
  int64_t launches = 0;
  std::vector<float *> saved;     // swsph_save_state copies
  // Chunk tables, one per chunk capacity in use, and the capacity used by
  // each loop (density, gradient, force).
  struct ChunkTable {
    int C = 0;
    int nchunks = 0;
    bool valid = false;    // built for the current particles
    bool stale = false;    // bounds out of date (refreshed on next use)
    int cap_chunks = 0;    // device capacity (chunks, cells)
    int cap_cells = 0;
    std::vector<int> cell_chunk_h;  // [ncells + 1]
    int *d_start = nullptr, *d_cell_chunk = nullptr, *d_chunk_cell = nullptr;
    double *d_bbox = nullptr;
    float *d_hg2max = nullptr;
    // boxes of the pre-test records (the tables of the sparse tiles)
    float4 *d_fbox = nullptr;
    // positions the boxes were formed from (engine pos_gen; 0: unknown)
    uint64_t pos_gen = 0;
  };
  std::vector<ChunkTable> tables;
  // Generation of the particle positions (bumped by every binning, drift
  // and grid change) and the one the pre-test records were formed from.
  uint64_t pos_gen = 1, posrec_gen = 0;
  int loop_c[3] = {16, 16, 16};   // tile (i-chunk) capacity CI per loop
  int loop_cj[3] = {16, 16, 16};  // candidate chunk capacity CJ (work queue)
  unsigned int *d_hg2_all = nullptr;  // max (gamma h)^2 over all chunks
  unsigned int *h_hg2_all = nullptr;
  // Neighbour lists of the batched pass (allocated on first use).
  bool use_lists = false;
  float list_skin = 1.2f;  // list radius / search radius
  int list_slots = 48;     // slots per stream
  NbrList nl{};
  size_t nl_elems = 0;
  int64_t list_rebuilt = 0;   // particles re-listed in the last pass
  int64_t list_fallback = 0;  // gradient fall-backs in the last pass
  // Schedule of the batched loops: 0 one wave per chunk over all chunks,
  // 1 one wave per tile of a device-built tile queue, 2 persistent waves
  // draining that queue.
  int schedule = 1;
  int queue_waves = 0;  // 0: one full-occupancy wave set per CU
  TileScratch ts{};
  size_t tile_cap = 0;  // entries of ts.tile_start / ts.tile_cell
  int tiles_c = 0;  // capacity of the all-active tile table in ts, 0: none
  // Steps with few active particles per cell (fewer than sparse_per_cell on
  // average) run their pass with tiles of sparse_ci against chunks of
  // sparse_cj in all three loops; 0 disables.
  int sparse_ci = 4, sparse_cj = 8, sparse_per_cell = 16;
  // the ghost / extra ghost / force records / end force of a pass sweep the
  // work queue's list instead of every slot (set by engine_step when under
  // half of the particles are active; an all-active pass sweeps slots,
  // 1 % faster there)
  bool sweep_list = false;
  // Device capacity (particles, cells): buffers are reused by later
  // swsph_set_particles calls that fit, without reallocation.
  size_t cap_n = 0;
  int cap_cells = 0;
  bool saved_valid = false;  // swsph_save_state holds this particle set
  bool need_particles = false;  // grid changed since the last upload
  // Graph replay of the batched pass: the density phase (with a fixed
  // number of re-runs) and the gradient-to-end-force phase, captured for
  // the configuration generation `graph_gen`.
  uint64_t gen = 1;  // bumped by every change of particles or settings
  uint64_t graph_gen = 0;
  int graph_reruns = 0;
  // redo count after each captured ghost round (device, pinned readback),
  // and how many replays in a row needed fewer rounds than captured
  static constexpr int kHistRounds = 64;
  unsigned long long *d_hist = nullptr, *h_hist = nullptr;
  int shrink_votes = 0;
  int last_rounds = 4;  // ghost rounds of the last pass
  bool capturing = false;
  hipGraphExec_t g_density = nullptr, g_finish = nullptr;
  std::vector<float *> tune_saved;  // pass state kept by the autotuner
  size_t saved_cap = 0, tune_cap = 0;  // particles each copy set holds
  // Device binning (sph_step.hip): key / value buffers of the radix sort,
  // its temporary storage, and a staging buffer of 3 doubles per particle
  // for uploads in the caller's order and for permutations.
  unsigned long long *d_key[2] = {nullptr, nullptr};
  int *d_val[2] = {nullptr, nullptr};
  void *d_sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  // In-cell order by time bin, then Morton (swsph_set_bin_order), and the
  // key layout of the current binning.
  bool bin_order = true;
  int key_bin_bits = 0;
  // record slot of each field slot and back, with the time-bin order (set
  // into d.cand / d.soa only while the binning uses that order)
  int *d_cand_buf = nullptr, *d_soa_buf = nullptr;
  unsigned int *d_moved = nullptr;  // set when the two orders differ
  double *d_stage = nullptr;
  // Time integration: parameters, the integer time line, and the largest
  // distance by which a particle lies outside its cell since the last
  // binning (float bits on the device; host copy).
  swsph_integration integ{0.1f, 0.f, 0.f, 0.0, 1e30, 0};
  int64_t ti_current = 0, ti_end_min = 0;
  bool started = false;
  unsigned int *d_dmax = nullptr;
  unsigned long long *d_tiend = nullptr;  // [1] ti_end_min
  // task bundles: device copies of the merged items and neighbour images
  BundleItem *d_bitems = nullptr;
  BundleNbr *d_bnbrs = nullptr;
  size_t cap_bitems = 0, cap_bnbrs = 0;
  unsigned int *d_err = nullptr;          // [3] time-step errors, active,
                                          // active count before a step
  unsigned int *h_small = nullptr;        // pinned readback: dmax, hg2
  unsigned long long *h_tiend = nullptr;  // pinned [1]
  unsigned int *h_err = nullptr;          // pinned [3]
  int *h_ntiles = nullptr;                // pinned [1]
  float dmax = 0.f;
  // The neighbour records of every particle match its fields (the passes
  // then rebuild only those of the particles they update), and the
  // generation of the fields a saved state was taken from.
  bool recs_ok = false;
  uint64_t fields_epoch = 1, saved_epoch = 0;
  // After a restore of the pass state, the records of the particles active
  // in the last pass describe the overwritten values: valid only if the
  // next pass updates the same particles (kNoMab: no such restore pending).
  static constexpr int kNoMab = -1000;
  int stale_mab = kNoMab;
  int last_pass_mab = kNoMab;
  // Device the engine was created on, and set while a pass runs on the
  // fixed scratch of swsph_pass_batched_noalloc: any path that would
  // allocate device memory fails instead.
  int device = 0;
  bool no_alloc = false;
  // Tiles of the all-active work queue per tile capacity, counted before a
  // graph capture (the captured grids are sized by them).
  std::map<int, int> cap_tiles;
  // Bundles of cell tasks (swsph_*_cells): the cells' mask on the device,
  // the uploaded cell list, and whether the mask is in force; the cells
  // holding active particles (swsph_sched_pass).
  unsigned char *d_cmask = nullptr, *d_cact = nullptr;
  int *d_clist = nullptr;
  int cmask_cap = 0;
  bool cmask_on = false;
  // swsph_engine_step runs its pass through the task graph (1) or the work
  // queue (0)
  int step_sched = 0;
  // Whole passes regrid the top level when the kernel support outgrows a
  // cell (swsph_set_regrid); regrids done since creation.
  bool auto_regrid = true;
  int64_t regrids = 0;
  // Subsets of the caller's particles (swsph_do*_subset_density): flags per
  // slot, the device slot of every caller index (valid for the binning
  // generation inv_gen), the uploaded indices, and an error counter;
  // allocated on first use.  perm_gen counts binnings.
  int *d_subflag = nullptr, *d_inv = nullptr, *d_subind = nullptr;
  unsigned int *d_suberr = nullptr;
  size_t sub_cap = 0, subind_cap = 0;
  uint64_t perm_gen = 1, inv_gen = 0;
};

namespace {
// The (io) state of a pass, in the order saved by swsph_save_state.
std::vector<float *> io_fields(DevParts &d) {
  return {d.h, d.rho, d.P, d.c, d.f, d.bal, d.alpha, d.adiff, d.divv_prev};
}
// Frees the swsph_save_state copies and the autotuner's copies: both are
// sized for the particle capacity and go whenever the capacity changes.
void free_saved(swsph_engine *e) {
  for (float *p : e->saved)
    if (p) (void)hipFree(p);
  e->saved.clear();
  for (float *p : e->tune_saved)
    if (p) (void)hipFree(p);
  e->tune_saved.clear();
  e->saved_cap = e->tune_cap = 0;
}
}  // namespace

namespace {

// Component error bound of the FP32 pre-test: 2^-19 of 1.5 widths of the
// widest cell, four times the worst case of the few float roundings that
// separate the pre-test distance from the contract's, for values bounded by
// two cell widths plus twice the distance d by which a drifted particle may
// lie outside its cell (valid for d <= w / 2, which the stencil condition
// H + 2 d <= w guarantees).
float pos_eps(const swsph_engine *e) {
  double w = 0.0;
  for (int k = 0; k < 3; k++) w = std::max(w, e->grid.dim[k] / e->grid.nc[k]);
  return (float)(1.5 * w * 0x1p-19);
}

// The records of some particle no longer match its fields: the next pass
// rebuilds all of them (captured graphs are re-captured).
void invalidate_records(swsph_engine *e) {
  if (e->recs_ok) e->gen++;
  e->recs_ok = false;
}

// Before a pass (never inside a graph capture): rebuild every particle's
// records if some do not match its fields.
int ensure_records(swsph_engine *e) {
  if (e->stale_mab != swsph_engine::kNoMab &&
      e->props.max_active_bin != e->stale_mab)
    invalidate_records(e);
  e->stale_mab = swsph_engine::kNoMab;
  if (e->recs_ok) return 0;
  for (LoopKind k : {LoopKind::Density, LoopKind::Gradient, LoopKind::Force})
    HIPCHK(sphk::launch_build_records(k, e->d, 0, (int)e->n, e->st));
  e->launches += 3;
  e->recs_ok = true;
  return 0;
}

LoopArgs loop_args(const swsph_engine *e) {
  LoopArgs la;
  la.grid = e->grid;
  la.cell_start = e->d_cell_start;
  la.flag = e->d.redo;
  la.max_active_bin = e->props.max_active_bin;
  la.pos_eps = pos_eps(e);
  la.cmask = e->cmask_on ? e->d_cmask : nullptr;
  la.fresh = 0;
  return la;
}

// The batched loops (tiles, stencil grid, list walks) give each particle
// its whole stencil in one wave: the sums start from their initial values
// in registers (L::fresh_acc) and the resets before them are skipped.
LoopArgs batched_args(const swsph_engine *e, LoopKind) {
  LoopArgs la = loop_args(e);
  la.fresh = 1;
  return la;
}

GhostArgs ghost_args(const swsph_engine *e) {
  GhostArgs g;
  const float eta = e->props.eta;
  g.eta_dim = eta * eta * eta;
  g.eps = e->props.h_tolerance;
  g.h_max = e->props.h_max;
  g.h_min = e->props.h_min;
  g.max_active_bin = e->props.max_active_bin;
  g.time_base = e->props.time_base;
  g.alpha_max = e->props.alpha_max;
  g.alpha_min = e->props.alpha_min;
  g.visc_length = e->props.visc_length;
  g.diff_beta = e->props.diff_beta;
  g.diff_alpha_max = e->props.diff_alpha_max;
  g.diff_alpha_min = e->props.diff_alpha_min;
  g.fresh_force = 0;
  g.fresh_grad = 0;
  g.lists = e->use_lists ? 1 : 0;
  g.grad_rec = 0;
  g.force_rec = 0;
  return g;
}

void free_dev(swsph_engine *e) {
  void *ptrs[] = {e->d.px, e->d.py, e->d.pz, e->d.vx, e->d.vy, e->d.vz,
                  e->d.m, e->d.u, e->d.h, e->d.tb, e->d.rho, e->d.rho_dh,
                  e->d.wc, e->d.wc_dh, e->d.divv, e->d.rotx, e->d.roty,
                  e->d.rotz, e->d.nd, e->d.left, e->d.right, e->d.redo,
                  e->d.P, e->d.c, e->d.f, e->d.bal, e->d.vsig, e->d.lapu,
                  e->d.avmax, e->d.ng, e->d.alpha, e->d.adiff, e->d.divv_prev,
                  e->d.divv_dt, e->d.ax, e->d.ay, e->d.az, e->d.udt, e->d.hdt,
                  e->d.nf, e->d.mintb, e->d.rd, e->d.rg, e->d.rf, e->d.pos,
                  e->d.r2b, e->d.lfb, e->d.sel, e->d.cell, e->nl.idx,
                  e->nl.cnt, e->ts.flag, e->ts.scan, e->ts.blk, e->ts.cnt,
                  e->ts.cscan, e->ts.cblk, e->ts.alist, e->ts.tile_start,
                  e->ts.tile_cell, e->ts.ntiles, e->ts.heads,
                  e->d_cell_start, e->d_tiles, e->d.vfx, e->d.vfy, e->d.vfz,
                  e->d.ufull, e->d.perm, e->d_key[0], e->d_key[1],
                  e->d_val[0], e->d_val[1], e->d_sort_tmp, e->d_stage,
                  e->d_subflag, e->d_inv, e->d_subind, e->d_suberr,
                  e->d_cmask, e->d_cact, e->d_clist, e->d_cand_buf,
                  e->d_soa_buf, e->d_moved};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  e->d_cand_buf = e->d_soa_buf = nullptr;
  e->d_moved = nullptr;
  e->d_subflag = e->d_inv = e->d_subind = nullptr;
  e->d_suberr = nullptr;
  e->sub_cap = e->subind_cap = 0;
  e->inv_gen = 0;
  e->d_cmask = e->d_cact = nullptr;
  e->d_clist = nullptr;
  e->cmask_cap = 0;
  e->cmask_on = false;
  e->d = DevParts{};
  e->ts = TileScratch{};
  e->nl.idx = nullptr;
  e->nl.cnt = nullptr;
  e->nl_elems = 0;
  e->d_cell_start = nullptr;
  e->d_tiles = nullptr;
  e->d_key[0] = e->d_key[1] = nullptr;
  e->d_val[0] = e->d_val[1] = nullptr;
  e->d_sort_tmp = nullptr;
  e->sort_tmp_bytes = 0;
  e->d_stage = nullptr;
  e->cap_n = 0;
  e->cap_cells = 0;
  free_saved(e);
  e->saved_valid = false;
  for (auto &t : e->tables) {
    void *tp[] = {t.d_start, t.d_cell_chunk, t.d_chunk_cell, t.d_bbox,
                  t.d_hg2max, t.d_fbox};
    for (void *p : tp)
      if (p) (void)hipFree(p);
  }
  e->tables.clear();
}

// Chunk table for capacity C: runs of <= C consecutive particles of a cell.
// Built on first use; boxes and radii are filled by refresh_bounds.
int refresh_bounds(swsph_engine *e, int c_lo, int c_hi, bool pos = true,
                   const swsph_engine::ChunkTable *extra = nullptr);

swsph_engine::ChunkTable *chunk_table(swsph_engine *e, int C, int *rc) {
  *rc = 0;
  swsph_engine::ChunkTable *tp = nullptr;
  for (auto &t : e->tables)
    if (t.C == C) tp = &t;
  if (tp && tp->valid) {
    if (tp->stale) {  // positions or radii changed while it was not in use
      const hipError_t err = sphk::launch_chunk_bounds(
          e->d, tp->d_start, 0, tp->nchunks, tp->C, tp->d_bbox, tp->d_hg2max,
          e->d_hg2_all, e->st, e->grid, pos_eps(e), tp->d_fbox);
      if (err != hipSuccess) {
        *rc = fail(2, "chunk bounds: %s", hipGetErrorString(err));
        return nullptr;
      }
      e->launches++;
      tp->stale = false;
      tp->pos_gen = e->posrec_gen;
    }
    return tp;
  }
  if (C != 8 && C != 16 && C != 32 && C != 64) {
    *rc = fail(1, "chunk capacity %d not in {8, 16, 32, 64}", C);
    return nullptr;
  }
  if (!tp) {
    if (e->tables.size() == e->tables.capacity()) {
      *rc = fail(2, "chunk table limit reached");
      return nullptr;
    }
    e->tables.emplace_back();
    tp = &e->tables.back();
    tp->C = C;
  }
  swsph_engine::ChunkTable &t = *tp;
  // first chunk of every cell on the host (ncells + 1 entries); the chunks'
  // starts and cells are filled on the device
  t.cell_chunk_h.assign(e->ncells + 1, 0);
  long long nch = 0;
  for (int c = 0; c < e->ncells; c++) {
    t.cell_chunk_h[c] = (int)nch;
    nch += (e->cell_start[c + 1] - e->cell_start[c] + C - 1) / C;
  }
  t.nchunks = (int)nch;
  t.cell_chunk_h[e->ncells] = t.nchunks;
  // (re)allocate only when the capacity does not suffice
  if (t.nchunks > t.cap_chunks || e->ncells > t.cap_cells) {
    if (e->no_alloc) {
      t.valid = false;
      *rc = fail(1, "the chunk table of capacity %d needs device memory: run "
                 "swsph_pass_batched once with these particles and shapes "
                 "before the allocation-free pass", C);
      return nullptr;
    }
    void *tp2[] = {t.d_start, t.d_cell_chunk, t.d_chunk_cell, t.d_bbox,
                   t.d_hg2max, t.d_fbox};
    for (void *q : tp2)
      if (q) (void)hipFree(q);
    t.d_start = t.d_cell_chunk = t.d_chunk_cell = nullptr;
    t.d_bbox = nullptr;
    t.d_hg2max = nullptr;
    t.d_fbox = nullptr;
    const size_t nk = (size_t)std::max(1, t.nchunks);
    if (dalloc(&t.d_start, nk + 1) != hipSuccess ||
        dalloc(&t.d_cell_chunk, (size_t)e->ncells + 1) != hipSuccess ||
        dalloc(&t.d_chunk_cell, nk) != hipSuccess ||
        dalloc(&t.d_bbox, 6 * nk) != hipSuccess ||
        dalloc(&t.d_hg2max, nk) != hipSuccess ||
        ((C == kSparseChunk || C == kSparseSub) &&
         dalloc(&t.d_fbox, 2 * nk) != hipSuccess)) {
      *rc = fail(2, "allocation of chunk table failed");
      return nullptr;
    }
    t.cap_chunks = t.nchunks;
    t.cap_cells = e->ncells;
  }
  if (hipMemcpyAsync(t.d_cell_chunk, t.cell_chunk_h.data(),
                     sizeof(int) * t.cell_chunk_h.size(),
                     hipMemcpyHostToDevice, e->st) != hipSuccess ||
      sphk::launch_chunk_fill(e->d_cell_start, t.d_cell_chunk, e->ncells, C,
                              t.nchunks, (int)e->n, t.d_start, t.d_chunk_cell,
                              e->st) != hipSuccess) {
    *rc = fail(2, "chunk table fill failed");
    return nullptr;
  }
  e->launches++;
  // the host array is rewritten by the next build of this table
  if (hipStreamSynchronize(e->st) != hipSuccess) {
    *rc = fail(2, "chunk table upload failed");
    return nullptr;
  }
  t.valid = true;
  t.stale = false;
  if ((*rc = refresh_bounds(e, 0, e->ncells, true, &t))) return nullptr;
  return tp;
}

ChunkArgs chunk_args(const swsph_engine::ChunkTable *t) {
  ChunkArgs ca;
  ca.chunk_start = t->d_start;
  ca.cell_chunk = t->d_cell_chunk;
  ca.chunk_cell = t->d_chunk_cell;
  ca.bbox = t->d_bbox;
  ca.hg2max = t->d_hg2max;
  ca.fbox = t->d_fbox;
  ca.nchunks = t->nchunks;
  ca.C = t->C;
  return ca;
}

// Capacity of the chunk table whose chunks serve as i-chunks (grid
// schedule, per-task entry points, task bundles): the tile capacity, or for
// the narrow tiles of the work queue (fewer than 8) their candidate chunks.
int chunk_c(const swsph_engine *e, int k) {
  return e->loop_c[k] >= 8 ? e->loop_c[k] : e->loop_cj[k];
}

// Chunk capacity of the table a loop reads: CJ in the work-queue schedules,
// else the i-chunk capacity.
int table_c(const swsph_engine *e, int k) {
  return e->schedule >= 1 ? e->loop_cj[k] : chunk_c(e, k);
}

// The sparse shape (tiles of 4, sub-chunks of 8) of the work queue: its
// loops also read the chunks of 32 of the same cells.
bool sparse_loop(const swsph_engine *e, int k) {
  return e->schedule >= 1 && e->loop_c[k] == kMinTile &&
         e->loop_cj[k] == kSparseSub;
}
// Capacity of the second table a loop reads (0: none).
int table2_c(const swsph_engine *e, int k) {
  return sparse_loop(e, k) ? kSparseChunk : 0;
}
// Builds (or refreshes) the tables loop k reads.
int ensure_tables(swsph_engine *e, int k) {
  int rc = 0;
  if (!chunk_table(e, table_c(e, k), &rc)) return rc;
  if (table2_c(e, k) && !chunk_table(e, table2_c(e, k), &rc)) return rc;
  return 0;
}

// Tables no loop reads any more are dropped from the per-pass refresh
// (rebuilt on their next use).
void prune_tables(swsph_engine *e) {
  for (auto &t : e->tables) {
    bool used = false;
    for (int k = 0; k < 3; k++)
      used = used || t.C == table_c(e, k) || t.C == chunk_c(e, k) ||
             t.C == table2_c(e, k);
    if (!used) t.valid = false;
  }
}

// Candidate positions / radii (unless `pos` is false: already written) and
// chunk bounds of cells [c_lo, c_hi), after positions or smoothing lengths
// changed.  Only the tables the loops read now (and `extra`) are refreshed;
// the others are marked stale and refreshed whole on their next use.
int refresh_bounds(swsph_engine *e, int c_lo, int c_hi, bool pos,
                   const swsph_engine::ChunkTable *extra) {
  const int lo = e->cell_start[c_lo], hi = e->cell_start[c_hi];
  const bool all = c_lo == 0 && c_hi == e->ncells;
  if (all) HIPCHK(sphk::launch_zero_u32(e->d_hg2_all, 1, e->st));
  // Positions unchanged since the records and the boxes in use were formed
  // (a pass moves no particle; only the ghost's radii change): radii only.
  bool radius = all && pos && e->posrec_gen == e->pos_gen;
  for (auto &t : e->tables) {
    if (!radius) break;
    bool u = &t == extra;
    for (int k = 0; k < 3; k++)
      u = u || t.C == table_c(e, k) || t.C == table2_c(e, k);
    if (u && t.valid && !t.stale && t.pos_gen != e->pos_gen) radius = false;
  }
  bool records_done = !pos;
  if (pos && !radius) {
    HIPCHK(sphk::launch_build_pos(e->d, e->grid, pos_eps(e), lo, hi, e->st));
    e->launches++;
    e->posrec_gen = all ? e->pos_gen : 0;
    records_done = true;
  }
  auto used = [&](const swsph_engine::ChunkTable &t) {
    bool u = &t == extra;
    for (int k = 0; k < 3; k++)
      u = u || t.C == table_c(e, k) || t.C == table2_c(e, k);
    return u;
  };
  auto refresh = [&](swsph_engine::ChunkTable &t) -> int {
    if (radius) {
      HIPCHK(sphk::launch_radius_bounds(
          e->d, t.d_start, 0, t.nchunks, t.C, t.d_hg2max, e->d_hg2_all,
          pos_eps(e), t.d_fbox, !records_done, e->st));
      records_done = true;
    } else {
      HIPCHK(sphk::launch_chunk_bounds(e->d, t.d_start, t.cell_chunk_h[c_lo],
                                       t.cell_chunk_h[c_hi], t.C, t.d_bbox,
                                       t.d_hg2max, e->d_hg2_all, e->st,
                                       e->grid, pos_eps(e), t.d_fbox));
      t.pos_gen = all && e->posrec_gen == e->pos_gen ? e->pos_gen : 0;
    }
    e->launches++;
    return 0;
  };
  bool any = false;
  for (auto &t : e->tables) {
    if (!t.valid || t.stale || !used(t)) continue;
    if (int rc = refresh(t)) return rc;
    any = true;
  }
  for (auto &t : e->tables) {
    if (!t.valid || t.stale || used(t)) continue;
    if (any) {
      t.stale = true;
    } else {
      // no table in use is built yet: this one carries the largest radius
      if (int rc = refresh(t)) return rc;
      any = true;
    }
  }
  if (!records_done) {  // no table took the records along
    HIPCHK(sphk::launch_build_pos(e->d, e->grid, pos_eps(e), lo, hi, e->st));
    e->launches++;
  }
  return 0;
}

// Device buffers for up to n particles and e->ncells cells.
int alloc_dev(swsph_engine *e, size_t n) {
  DevParts &d = e->d;
  HIPCHK(dalloc(&e->d_cell_start, (size_t)e->ncells + 1));
  HIPCHK(dalloc(&d.px, n));
  HIPCHK(dalloc(&d.py, n));
  HIPCHK(dalloc(&d.pz, n));
  float **f32[] = {&d.vx, &d.vy, &d.vz, &d.m, &d.u, &d.h, &d.rho, &d.rho_dh,
                   &d.wc, &d.wc_dh, &d.divv, &d.rotx, &d.roty, &d.rotz,
                   &d.left, &d.right, &d.P, &d.c, &d.f, &d.bal, &d.vsig,
                   &d.lapu, &d.avmax, &d.alpha, &d.adiff, &d.divv_prev,
                   &d.divv_dt, &d.ax, &d.ay, &d.az, &d.udt, &d.hdt,
                   &d.vfx, &d.vfy, &d.vfz, &d.ufull};
  for (float **p : f32) HIPCHK(dalloc(p, n));
  HIPCHK(dalloc(&d.r2b, n));
  int **i32[] = {&d.nd, &d.redo, &d.ng, &d.nf, &d.lfb, &d.sel, &d.cell,
                 &d.perm};
  for (int **p : i32) HIPCHK(dalloc(p, n));
  HIPCHK(dalloc(&d.tb, n));
  HIPCHK(dalloc(&d.mintb, n));
  HIPCHK(dalloc(&d.rd, n));
  HIPCHK(dalloc(&d.rg, n));
  HIPCHK(dalloc(&d.rf, n));
  HIPCHK(dalloc(&d.pos, n));
  HIPCHK(dalloc(&e->d_cand_buf, n));
  HIPCHK(dalloc(&e->d_soa_buf, n));
  HIPCHK(dalloc(&e->d_moved, 1));
  {
    TileScratch &t = e->ts;
    const size_t nc = (size_t)e->ncells;
    HIPCHK(dalloc(&t.flag, n));
    HIPCHK(dalloc(&t.scan, n + 1));
    HIPCHK(dalloc(&t.blk, n / 1024 + 2));
    HIPCHK(dalloc(&t.cnt, nc));
    HIPCHK(dalloc(&t.cscan, nc + 1));
    HIPCHK(dalloc(&t.cblk, nc / 1024 + 2));
    HIPCHK(dalloc(&t.alist, n));
    // a cell of k particles makes ceil(k / C) tiles: at most n / C + nc
    // tiles for tiles of C >= kMinTile particles, plus the end marker
    e->tile_cap = n / kMinTile + nc + 1;
    HIPCHK(dalloc(&t.tile_start, e->tile_cap));
    HIPCHK(dalloc(&t.tile_cell, e->tile_cap));
    HIPCHK(dalloc(&t.ntiles, 1));
    HIPCHK(dalloc(&t.heads, 8 * 64));
  }
  // binning buffers
  HIPCHK(dalloc(&e->d_key[0], n));
  HIPCHK(dalloc(&e->d_key[1], n));
  HIPCHK(dalloc(&e->d_val[0], n));
  HIPCHK(dalloc(&e->d_val[1], n));
  HIPCHK(dalloc(&e->d_stage, 3 * n));
  size_t bytes = 0;
  HIPCHK(sphk::sort_pairs(nullptr, bytes, e->d_key[0], e->d_key[1],
                          e->d_val[0], e->d_val[1], (int)n, 64, e->st));
  HIPCHK(hipMalloc(&e->d_sort_tmp, std::max<size_t>(bytes, 16)));
  e->sort_tmp_bytes = std::max<size_t>(bytes, 16);
  e->cap_n = n;
  e->cap_cells = e->ncells;
  return 0;
}

// Zero the particle fields of a new particle set of n (so that fields never
// written for inactive particles read as zero) and forget its lists.
int reset_dev(swsph_engine *e, size_t n) {
  DevParts &d = e->d;
  float *f32[] = {d.vx, d.vy, d.vz, d.m, d.u, d.h, d.rho, d.rho_dh,
                  d.wc, d.wc_dh, d.divv, d.rotx, d.roty, d.rotz,
                  d.left, d.right, d.P, d.c, d.f, d.bal, d.vsig,
                  d.lapu, d.avmax, d.alpha, d.adiff, d.divv_prev,
                  d.divv_dt, d.ax, d.ay, d.az, d.udt, d.hdt,
                  d.vfx, d.vfy, d.vfz, d.ufull};
  int *i32[] = {d.nd, d.redo, d.ng, d.nf, d.lfb, d.sel, d.cell, d.perm};
  const size_t nn = n ? n : 1;
  for (float *p : f32) HIPCHK(hipMemsetAsync(p, 0, sizeof(float) * nn, e->st));
  for (int *p : i32) HIPCHK(hipMemsetAsync(p, 0, sizeof(int) * nn, e->st));
  HIPCHK(hipMemsetAsync(d.mintb, 0, nn, e->st));
  HIPCHK(sphk::launch_fill_f32(d.r2b, -1.f, (int)nn, e->st));  // no lists yet
  HIPCHK(hipMemsetAsync(e->d_dmax, 0, sizeof(unsigned int), e->st));
  e->dmax = 0.f;
  for (auto &t : e->tables) t.valid = false;
  e->saved_valid = false;
  e->tiles_c = 0;
  return 0;
}

int check_props(const swsph_props *p) {
  for (int k = 0; k < 3; k++) {
    if (p->ncell[k] < 3)
      return fail(1, "need >= 3 top-level cells per axis (got %d)", p->ncell[k]);
    if (!(p->dim[k] > 0.0)) return fail(1, "box size must be positive");
  }
  if (p->max_iter <= 0) return fail(1, "max_iter must be positive");
  if (!(p->eta > 0.f)) return fail(1, "eta must be positive");
  return 0;
}

bool adjacent(const swsph_engine *e, int ci, int cj, double t[3]) {
  const int nc[3] = {e->grid.nc[0], e->grid.nc[1], e->grid.nc[2]};
  const int a[3] = {ci / (nc[1] * nc[2]), (ci / nc[2]) % nc[1], ci % nc[2]};
  const int b[3] = {cj / (nc[1] * nc[2]), (cj / nc[2]) % nc[1], cj % nc[2]};
  for (int k = 0; k < 3; k++) {
    int dd = b[k] - a[k];
    t[k] = 0.0;
    if (dd > nc[k] / 2) {
      dd -= nc[k];
      t[k] = -e->grid.dim[k];
    } else if (dd < -(nc[k] / 2)) {
      dd += nc[k];
      t[k] = e->grid.dim[k];
    }
    if (dd < -1 || dd > 1) return false;
  }
  return true;
}

int cell_ok(const swsph_engine *e, int c) {
  if (e->n == 0 && e->cell_start.empty()) return fail(1, "no particles set");
  if (e->need_particles)
    return fail(1, "the grid changed: call swsph_set_particles first");
  if (c < 0 || c >= e->ncells) return fail(1, "cell index %d out of range", c);
  return 0;
}

hipEvent_t next_event(swsph_engine *e) {
  if (e->ev_used == e->ev_pool.size()) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    e->ev_pool.push_back(ev);
  }
  return e->ev_pool[e->ev_used++];
}

// Timing events around the launches of one loop kind, when enabled.
int time_begin(swsph_engine *e, hipEvent_t *a) {
  *a = nullptr;
  if (!e->timing || e->capturing) return 0;
  *a = next_event(e);
  if (!*a) return fail(2, "hipEventCreate failed");
  HIPCHK(hipEventRecord(*a, e->st));
  return 0;
}

int time_end(swsph_engine *e, LoopKind k, hipEvent_t a) {
  if (!e->timing || e->capturing) return 0;
  hipEvent_t b = next_event(e);
  if (!b) return fail(2, "hipEventCreate failed");
  HIPCHK(hipEventRecord(b, e->st));
  e->ev_pending.push_back({(int)k, {a, b}});
  return 0;
}

// Launch a batched loop; a subset loop takes the particles with flag[i].
int stencil(swsph_engine *e, LoopKind k, bool subset,
            const int *flag = nullptr) {
  int rc = 0;
  const swsph_engine::ChunkTable *t = chunk_table(e, chunk_c(e, (int)k), &rc);
  if (!t) return rc;
  hipEvent_t a;
  if ((rc = time_begin(e, &a))) return rc;
  LoopArgs la = batched_args(e, k);
  if (flag) la.flag = flag;
  HIPCHK(sphk::launch_stencil(k, subset, e->d, la, chunk_args(t), e->st));
  e->launches++;
  return time_end(e, k, a);
}

TileArgs tile_args(const swsph_engine *e) {
  TileArgs ta;
  ta.alist = e->ts.alist;
  ta.tile_start = e->ts.tile_start;
  ta.tile_cell = e->ts.tile_cell;
  ta.ntiles = e->ts.ntiles;
  ta.heads = e->ts.heads;
  ta.ids = nullptr;
  ta.filt = nullptr;
  return ta;
}

// A batched loop through the work queue: tiles of the particles selected by
// `mode` (0 active, 1 active and flagged, 2 flagged without a usable
// neighbour list), drained by persistent waves; `build` rebuilds the lists
// of the selected particles while doing their density.
int queue_loop(swsph_engine *e, LoopKind k, int mode, const int *flag,
               bool build) {
  int rc = 0;
  const int C = build ? kListChunk : e->loop_c[(int)k];
  const int CJ = build ? kListChunk : e->loop_cj[(int)k];
  if ((size_t)e->n / C + (size_t)e->ncells + 1 > e->tile_cap)
    return fail(1, "tiles of %d particles exceed the tile table", C);
  const swsph_engine::ChunkTable *t = chunk_table(e, CJ, &rc);
  if (!t) return rc;
  const bool sparse = !build && sparse_loop(e, (int)k);
  const swsph_engine::ChunkTable *t1 =
      sparse ? chunk_table(e, kSparseChunk, &rc) : nullptr;
  if (sparse && !t1) return rc;
  const LoopArgs la = batched_args(e, k);
  TileArgs ta = tile_args(e);
  // The all-active table is kept for the later loops of the pass.  With
  // the narrowest tiles the flagged rounds (mode 1) take the subset of its
  // tiles that hold a flagged particle, with the others masked out, instead
  // of a new table (a build sweeps and scans every slot; this sweeps the
  // tiles).  Wider tiles rebuild: the flagged particles of a dense box are
  // scattered over nearly every wide tile, which would run mostly masked
  // (+20 % on the 128^3 box's re-run rounds).
  const bool subset = mode == 1 && e->tiles_c == C && C == kMinTile &&
                      !build && !la.cmask;
  if (subset) {
    ta.ids = e->ts.flag;    // free until the next build (tiles <= particles)
    ta.ntiles = e->ts.blk;  // the subset's count
    ta.filt = flag;
    HIPCHK(sphk::launch_select_tiles(e->ts, flag, e->ts.flag, e->ts.blk,
                                     e->st));
    if (e->schedule == 2 || e->capturing)
      HIPCHK(sphk::launch_zero_u32((unsigned int *)e->ts.heads, 8 * 64,
                                   e->st));
  } else if (!(mode == 0 && e->tiles_c == C)) {
    HIPCHK(sphk::launch_build_tiles(e->d, la, (int)e->n, e->ncells, C, mode,
                                    flag, e->ts, e->st));
    e->tiles_c = mode == 0 ? C : 0;
  } else if (e->schedule == 2) {
    HIPCHK(sphk::launch_zero_u32((unsigned int *)e->ts.heads, 8 * 64, e->st));
  }
  int max_tiles = (int)(e->n / C) + e->ncells + 1;
  bool persistent = e->schedule == 2, rest = false;
  if (!e->capturing && e->schedule == 1) {
    // Eager: one wave per tile of the device-side count (a grid for the
    // largest possible count would mostly be waves that exit at once).
    HIPCHK(hipMemcpyAsync(e->h_ntiles, ta.ntiles, sizeof(int),
                          hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    max_tiles = *e->h_ntiles;
  } else if (e->capturing && e->schedule == 1) {
    // Captured: the all-active tiles are counted before the capture (the
    // count holds for every replay of this configuration), with a small
    // grid behind for any tiles beyond it.  The re-run rounds' counts vary
    // but never exceed it (their particles are active ones): with wide
    // tiles, whose re-runs of a dense box hold most tiles again, they take
    // the same grid (surplus waves exit at once); narrow tiles' re-runs,
    // often a few tiles, are drained by the persistent grid.
    const auto it = e->cap_tiles.find(C);
    if (it != e->cap_tiles.end() && (mode == 0 || C > kMinTile)) {
      max_tiles = std::min(max_tiles, it->second);
      rest = true;
    } else if (mode != 0) {
      persistent = true;
    }
  }
  // the loop kernel alone is timed (the tile build is not)
  hipEvent_t a;
  if ((rc = time_begin(e, &a))) return rc;
  if ((max_tiles > 0 || rest) && sparse)
    HIPCHK(sphk::launch_sparse(k, e->d, la, chunk_args(t1), chunk_args(t), ta,
                               persistent, e->queue_waves, max_tiles, rest,
                               e->st));
  else if (max_tiles > 0 || rest)
    HIPCHK(sphk::launch_queue(k, build, e->d, la, chunk_args(t), ta, e->nl, C,
                              CJ, persistent, e->queue_waves, max_tiles, rest,
                              e->st));
  e->launches += 13;
  return time_end(e, k, a);
}

// Neighbour-list storage for the current particle count and slot count.
int ensure_lists(swsph_engine *e) {
  const size_t npad = std::max<size_t>(1, e->cap_n);
  const size_t elems = (size_t)e->list_slots * npad * kListStreams;
  if (e->nl.idx && e->nl_elems == elems) return 0;
  if (e->nl.idx) (void)hipFree(e->nl.idx);
  if (e->nl.cnt) (void)hipFree(e->nl.cnt);
  e->nl.idx = nullptr;
  e->nl.cnt = nullptr;
  e->nl_elems = 0;
  if (hipMalloc((void **)&e->nl.idx, sizeof(unsigned int) * elems) != hipSuccess ||
      hipMalloc((void **)&e->nl.cnt, sizeof(int) * npad * kListStreams) !=
          hipSuccess)
    return fail(2, "allocation of neighbour lists (%zu entries) failed", elems);
  e->nl_elems = elems;
  e->nl.Ks = e->list_slots;
  e->nl.npad = (int)npad;
  e->nl.lam2 = e->list_skin * e->list_skin;
  return 0;
}

// Density loop that (re)builds neighbour lists: over all active particles,
// or (subset) the ones flagged for another round whose list is too short.
int build_lists(swsph_engine *e, bool subset) {
  int rc = 0;
  const swsph_engine::ChunkTable *t = chunk_table(e, kListChunk, &rc);
  if (!t) return rc;
  hipEvent_t a;
  if ((rc = time_begin(e, &a))) return rc;
  HIPCHK(sphk::launch_build_lists(subset, e->d,
                                  batched_args(e, LoopKind::Density),
                                  chunk_args(t),
                                  e->nl, e->st));
  e->launches++;
  return time_end(e, LoopKind::Density, a);
}

// A density / gradient loop over the neighbour lists: all active particles,
// or the nsel particles of d.sel.
int list_loop(swsph_engine *e, LoopKind k, bool compact, int nsel) {
  int rc = 0;
  const swsph_engine::ChunkTable *t = chunk_table(e, kListChunk, &rc);
  if (!t) return rc;
  hipEvent_t a;
  if ((rc = time_begin(e, &a))) return rc;
  HIPCHK(sphk::launch_list_loop(k, compact, e->d, batched_args(e, k),
                                chunk_args(t),
                                e->nl, nsel, e->st));
  e->launches++;
  return time_end(e, k, a);
}

int collect_timing(swsph_engine *e) {
  if (!e->timing) return 0;
  HIPCHK(hipStreamSynchronize(e->st));
  for (auto &p : e->ev_pending) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    e->t_ms[p.first] += ms;
    e->t_launch[p.first] += 1;
  }
  e->ev_pending.clear();
  e->ev_used = 0;
  return 0;
}

int task(swsph_engine *e, LoopKind k, bool subset, int ci, int cj, bool both,
         const int *flag = nullptr) {
  int rc;
  if ((rc = cell_ok(e, ci)) || (rc = cell_ok(e, cj))) return rc;
  if ((rc = ensure_records(e))) return rc;
  TaskArgs ta{};
  double t[3];
  if (!adjacent(e, ci, cj, t))
    return fail(1, "cells %d and %d are not neighbours", ci, cj);
  const swsph_engine::ChunkTable *tb = chunk_table(e, chunk_c(e, (int)k), &rc);
  if (!tb) return rc;
  auto nch = [&](int c) { return tb->cell_chunk_h[c + 1] - tb->cell_chunk_h[c]; };
  ta.icell[0] = ci;
  ta.jcell[0] = cj;
  ta.ntiles[0] = nch(ci);
  for (int k2 = 0; k2 < 3; k2++) ta.t[k2] = t[k2];
  if (both && ci != cj) {
    ta.icell[1] = cj;
    ta.jcell[1] = ci;
    ta.ntiles[1] = nch(cj);
    for (int k2 = 0; k2 < 3; k2++) ta.t[3 + k2] = -t[k2];
  }
  if (ta.ntiles[0] + ta.ntiles[1] == 0) return 0;
  LoopArgs la = loop_args(e);
  if (flag) la.flag = flag;
  HIPCHK(sphk::launch_task(k, subset, e->d, la, chunk_args(tb), ta, e->st));
  e->launches++;
  return 0;
}

// The 13 "upper" neighbours of a cell: offsets whose first non-zero
// component is positive, so that every adjacent pair is visited once.
void upper_neighbours(const swsph_engine *e, int ci, std::vector<int> &out) {
  out.clear();
  const int *nc = e->grid.nc;
  const int a[3] = {ci / (nc[1] * nc[2]), (ci / nc[2]) % nc[1], ci % nc[2]};
  for (int o = 14; o < 27; o++) {
    const int ox = o / 9 - 1, oy = (o / 3) % 3 - 1, oz = o % 3 - 1;
    const int b[3] = {(a[0] + ox + nc[0]) % nc[0], (a[1] + oy + nc[1]) % nc[1],
                      (a[2] + oz + nc[2]) % nc[2]};
    out.push_back((b[0] * nc[1] + b[1]) * nc[2] + b[2]);
  }
}

void all_neighbours(const swsph_engine *e, int ci, std::vector<int> &out) {
  out.clear();
  const int *nc = e->grid.nc;
  const int a[3] = {ci / (nc[1] * nc[2]), (ci / nc[2]) % nc[1], ci % nc[2]};
  for (int o = 0; o < 27; o++) {
    if (o == 13) continue;
    const int ox = o / 9 - 1, oy = (o / 3) % 3 - 1, oz = o % 3 - 1;
    const int b[3] = {(a[0] + ox + nc[0]) % nc[0], (a[1] + oy + nc[1]) % nc[1],
                      (a[2] + oz + nc[2]) % nc[2]};
    out.push_back((b[0] * nc[1] + b[1]) * nc[2] + b[2]);
  }
}

// Ghost counters: [0] particles flagged for another round, [1] of those
// needing a new neighbour list, [2] converged without a usable list.
int read_redo(swsph_engine *e, int *count, int *relist = nullptr,
              int *nolist = nullptr) {
  HIPCHK(hipMemcpyAsync(e->h_redo, e->d_redo,
                        kGhostCounters * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  *count = (int)e->h_redo[0];
  if (relist) *relist = (int)e->h_redo[1];
  if (nolist) *nolist = (int)e->h_redo[2];
  int64_t work = 0;  // accumulated on the device, over the slots
  for (int k = 0; k < kGhostWorkSlots; k++) work += (int64_t)e->h_redo[3 + k];
  e->density_work = work;
  return 0;
}

// Device buffers of the caller-subset entry points: slot flags and the
// slot of every caller index for the current binning, and room for
// `count` uploaded indices.
int ensure_subset(swsph_engine *e, int count) {
  if (!e->d_subflag || e->sub_cap < e->cap_n) {
    for (void *p : {(void *)e->d_subflag, (void *)e->d_inv, (void *)e->d_suberr})
      if (p) (void)hipFree(p);
    e->d_subflag = e->d_inv = nullptr;
    e->d_suberr = nullptr;
    e->sub_cap = 0;
    e->inv_gen = 0;
    const size_t n = std::max<size_t>(1, e->cap_n);
    HIPCHK(dalloc(&e->d_subflag, n));
    HIPCHK(dalloc(&e->d_inv, n));
    HIPCHK(dalloc(&e->d_suberr, 1));
    e->sub_cap = n;
  }
  if ((size_t)count > e->subind_cap) {
    if (e->d_subind) (void)hipFree(e->d_subind);
    e->d_subind = nullptr;
    e->subind_cap = 0;
    HIPCHK(dalloc(&e->d_subind, (size_t)count));
    e->subind_cap = (size_t)count;
  }
  if (e->inv_gen != e->perm_gen) {
    HIPCHK(sphk::launch_invert_perm(e->d.perm, e->d_inv, (int)e->n, e->st));
    e->launches++;
    e->inv_gen = e->perm_gen;
  }
  return 0;
}

// runner_do_ghost (runner_ghost.c:1113) for share `off` of `nt` of cell ci's
// particles: rounds of the ghost kernel over the share, each followed by
// the subset re-runs of the share's flagged particles over every density
// interaction of the cell (runner_ghost.c:1548-1580).  The last share to
// finish leaves the cell's gradient records and chunk bounds current.
int ghost_cell(swsph_engine *e, int ci, int *rounds, int off = 0, int nt = 1) {
  const int lo = e->cell_start[ci], hi = e->cell_start[ci + 1];
  const GhostArgs ga = ghost_args(e);
  std::vector<int> nb;
  all_neighbours(e, ci, nb);
  int redo = 0;
  int it = 0;
  int rc = 0;
  // with shares, the re-runs take the flags of this share only (the other
  // shares' particles keep the flags hydro_init_part set until their own
  // ghost runs)
  const int *flag = nullptr;
  if (nt > 1) {
    if ((rc = ensure_subset(e, 0))) return rc;
    flag = e->d_subflag;
  }
  for (; it < e->props.max_iter; it++) {
    HIPCHK(sphk::launch_ghost(e->d, ga, lo, hi, e->d_redo, e->st, nullptr,
                              nullptr, off, nt));
    e->launches += 2;
    if ((rc = read_redo(e, &redo))) return rc;
    if (redo == 0) break;
    if (flag) {
      HIPCHK(sphk::launch_share_flags(e->d_subflag, e->d.redo, lo, hi, off,
                                      nt, e->st));
      e->launches++;
    }
    if ((rc = task(e, LoopKind::Density, true, ci, ci, false, flag))) return rc;
    for (int cj : nb)
      if ((rc = task(e, LoopKind::Density, true, ci, cj, false, flag)))
        return rc;
  }
  if (redo > 0)
    return fail(3, "Smoothing length failed to converge on %d particles", redo);
  if (rounds) *rounds = std::max(*rounds, it + 1);
  // The gradient loop reads rho, c, u, alpha of this cell's particles; the
  // candidate radii and chunk bounds follow the new smoothing lengths.
  HIPCHK(sphk::launch_build_records(LoopKind::Gradient, e->d, lo, hi, e->st));
  e->launches++;
  return refresh_bounds(e, ci, ci + 1);
}

// runner_do_extra_ghost on one cell, then the force records of its particles.
int extra_ghost_cell(swsph_engine *e, int c) {
  const int lo = e->cell_start[c], hi = e->cell_start[c + 1];
  HIPCHK(sphk::launch_extra_ghost(e->d, ghost_args(e), lo, hi, e->st));
  HIPCHK(sphk::launch_build_records(LoopKind::Force, e->d, lo, hi, e->st));
  e->launches += 2;
  return 0;
}

int count_into(swsph_engine *e, swsph_pass_stats *s) {
  HIPCHK(sphk::launch_count(e->d, (int)e->n, e->props.max_active_bin, e->d_cnt,
                            e->st));
  HIPCHK(hipMemcpyAsync(e->h_cnt, e->d_cnt, 3 * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  if (s) {
    s->n_density = (int64_t)e->h_cnt[0];
    s->n_gradient = (int64_t)e->h_cnt[1];
    s->n_force = (int64_t)e->h_cnt[2];
  }
  return 0;
}

// Kernel support H (largest over the chunk bounds) and drift distance d
// (largest distance of a particle outside its cell since the last binning)
// of the current particles.  Every neighbour lies in the 27-cell stencil
// while H + 2 d <= the cell width (the reference rebuilds its cells when
// that fails).
int read_fit(swsph_engine *e, double *H, double *dm) {
  HIPCHK(hipMemcpyAsync(e->h_small, e->d_hg2_all, sizeof(unsigned int),
                        hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipMemcpyAsync(e->h_small + 1, e->d_dmax, sizeof(unsigned int),
                        hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  float hg2, d;
  std::memcpy(&hg2, e->h_small, sizeof(float));
  std::memcpy(&d, e->h_small + 1, sizeof(float));
  e->dmax = d;
  *H = std::sqrt((double)hg2);
  *dm = (double)d;
  return 0;
}

bool fits(const swsph_engine *e, double H, double dm) {
  for (int k = 0; k < 3; k++)
    if (H + 2.0 * dm > e->grid.dim[k] / e->grid.nc[k]) return false;
  return true;
}

int check_fits(swsph_engine *e) {
  double H, dm;
  int rc = read_fit(e, &H, &dm);
  if (rc) return rc;
  if (!fits(e, H, dm))
    return fail(4,
                "kernel support %g plus twice the drift %g exceeds the "
                "top-level cell width %g: regrid with fewer cells",
                H, dm, e->grid.dim[0] / e->grid.nc[0]);
  return 0;
}

// After a pass: the largest kernel support met by the chunk bounds (and the
// drift since the last binning) must fit the stencil.
int check_support(swsph_engine *e) { return check_fits(e); }

int rebin(swsph_engine *e);

// A coarser top-level grid for the kernel support H, as space_regrid
// chooses it (src/space_regrid.c:118-126): floor(dim / (H space_stretch))
// cells per axis, space_stretch = 1.10 (src/space.h:74), never finer than
// the current grid; the particles are re-binned into it on the device.
int regrid(swsph_engine *e, double H) {
  int nc[3];
  bool same = true;
  for (int k = 0; k < 3; k++) {
    nc[k] = std::min(e->grid.nc[k],
                     (int)std::floor(e->grid.dim[k] / (H * 1.10)));
    if (nc[k] < 3)
      return fail(4, "kernel support %g needs fewer than 3 top-level cells "
                  "per axis of the box (%g)", H, e->grid.dim[k]);
    same = same && nc[k] == e->grid.nc[k];
  }
  if (same)
    return fail(4, "kernel support %g exceeds the top-level cell width %g",
                H, e->grid.dim[0] / e->grid.nc[0]);
  for (int k = 0; k < 3; k++) {
    e->grid.nc[k] = nc[k];
    e->props.ncell[k] = nc[k];
  }
  e->ncells = nc[0] * nc[1] * nc[2];
  for (auto &t : e->tables) t.valid = false;
  e->cap_tiles.clear();
  e->regrids++;
  return rebin(e);
}

// After a whole pass that ended with status 4 (kernel support beyond the
// 27-cell stencil): re-bin when the drift's slack alone broke it, else
// regrid (when allowed); the caller repeats the pass from the smoothing
// lengths found so far.
int make_fit(swsph_engine *e) {
  double H, dm;
  int rc = read_fit(e, &H, &dm);
  if (rc) return rc;
  if (dm > 0.0 && fits(e, H, 0.0)) return rebin(e);
  if (!e->auto_regrid || e->no_alloc)
    return fail(4, "kernel support %g plus twice the drift %g exceeds the "
                "top-level cell width %g", H, dm,
                e->grid.dim[0] / e->grid.nc[0]);
  return regrid(e, H);
}

// Runs a whole pass, re-binning or regridding and repeating it while the
// kernel support does not fit (at most a few times).
template <class F>
int fitting(swsph_engine *e, swsph_pass_stats *stats, F pass) {
  const int64_t r0 = e->regrids;
  int rc = pass();
  for (int k = 0; rc == 4 && k < 4; k++) {
    if ((rc = make_fit(e))) break;
    rc = pass();
  }
  if (stats) stats->regrids = e->regrids - r0;
  return rc;
}

// Keys, stable radix sort and cell bounds of the current n positions
// (component k of particle i at p_k[stride * i]); the sort's values are the
// positions' indices, written to vals_out in the new order.
// tb: the particles' time bins in the positions' index space (used when the
// engine orders cells by time bin).
int bin_sort(swsph_engine *e, const double *px, const double *py,
             const double *pz, int stride, int *vals_out, const int8_t *tb) {
  const int ni = (int)e->n;
  double inv_w[3];
  // iwidth as space_regrid sets it (space_regrid.c:221-222)
  for (int k = 0; k < 3; k++)
    inv_w[k] = 1.0 / (e->grid.dim[k] / e->grid.nc[k]);
  const bool by_bin = e->bin_order && tb;
  e->key_bin_bits = by_bin ? sphk::kBinBits : 0;
  HIPCHK(sphk::launch_bin_keys(px, py, pz, stride, e->grid, inv_w, ni,
                               e->d_key[0], e->d_val[0], e->st,
                               by_bin ? tb : nullptr));
  const int low = 30 + e->key_bin_bits;
  int bits = low;
  while ((1LL << (bits - low)) < (long long)e->ncells) bits++;
  size_t bytes = e->sort_tmp_bytes;
  HIPCHK(sphk::sort_pairs(e->d_sort_tmp, bytes, e->d_key[0], e->d_key[1],
                          e->d_val[0], vals_out, ni, bits, e->st));
  HIPCHK(sphk::launch_cell_bounds(e->d_key[1], ni, e->ncells,
                                  e->d_cell_start, e->d.cell, e->st,
                                  e->key_bin_bits));
  e->launches += 4;
  return 0;
}

// After a binning: host copy of the cell bounds, neighbour records of every
// particle (inactive particles keep theirs for the whole pass), candidate
// positions and the chunk tables of the loops.
int finish_binning(swsph_engine *e) {
  int rc = 0;
  e->pos_gen++;  // particles moved between slots
  e->d.cand = e->d.soa = nullptr;
  if (e->key_bin_bits && e->n > 0) {
    // Fields are in (cell, time bin, Morton) order; the records and chunk
    // tables take (cell, Morton) order: a stable sort of the field slots by
    // the key without the time bin gives the field slot of every record
    // slot, and its inverse the record slot of every field slot.
    const int ni = (int)e->n;
    double inv_w[3];
    for (int k = 0; k < 3; k++)
      inv_w[k] = 1.0 / (e->grid.dim[k] / e->grid.nc[k]);
    HIPCHK(sphk::launch_bin_keys(e->d.px, e->d.py, e->d.pz, 1, e->grid, inv_w,
                                 ni, e->d_key[0], e->d_val[0], e->st));
    int bits = 30;
    while ((1LL << (bits - 30)) < (long long)e->ncells) bits++;
    size_t bytes = e->sort_tmp_bytes;
    HIPCHK(sphk::sort_pairs(e->d_sort_tmp, bytes, e->d_key[0], e->d_key[1],
                            e->d_val[0], e->d_soa_buf, ni, bits, e->st));
    HIPCHK(sphk::launch_invert_perm(e->d_soa_buf, e->d_cand_buf, ni, e->st));
    HIPCHK(sphk::launch_perm_moved(e->d_soa_buf, ni, e->d_moved, e->st));
    e->launches += 4;
  }
  unsigned int moved = 0;
  if (e->key_bin_bits && e->n > 0)
    HIPCHK(hipMemcpyAsync(&moved, e->d_moved, sizeof(moved),
                          hipMemcpyDeviceToHost, e->st));
  e->cell_start.resize(e->ncells + 1);
  HIPCHK(hipMemcpyAsync(e->cell_start.data(), e->d_cell_start,
                        sizeof(int) * (e->ncells + 1), hipMemcpyDeviceToHost,
                        e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  if (moved) {  // otherwise every slot is its own (e.g. a single time bin)
    e->d.cand = e->d_cand_buf;
    e->d.soa = e->d_soa_buf;
  }
  for (LoopKind k : {LoopKind::Density, LoopKind::Gradient, LoopKind::Force})
    HIPCHK(sphk::launch_build_records(k, e->d, 0, (int)e->n, e->st));
  HIPCHK(sphk::launch_build_pos(e->d, e->grid, pos_eps(e), 0, (int)e->n,
                                e->st));
  e->recs_ok = true;
  e->stale_mab = swsph_engine::kNoMab;
  e->fields_epoch++;
  e->perm_gen++;
  for (auto &t : e->tables) t.valid = false;
  for (int k = 0; k < 3; k++)
    if ((rc = ensure_tables(e, k))) return rc;
  HIPCHK(hipStreamSynchronize(e->st));
  return 0;
}

// Re-binning of drifted particles (space_rebuild's particle sort for this
// path): positions wrapped into the box, sorted by (cell, Morton code) on
// the device, every per-particle field carried along.  Neighbour lists,
// saved states and captured graphs refer to the old order and are dropped.
int rebin(swsph_engine *e) {
  const int ni = (int)e->n;
  DevParts &d = e->d;
  int rc = 0;
  HIPCHK(sphk::launch_wrap(d, e->grid, ni, e->st));
  int *idx = e->d_val[1];
  if ((rc = bin_sort(e, d.px, d.py, d.pz, 1, idx, d.tb))) return rc;
  auto carry = [&](auto *field) -> int {
    using T = typename std::remove_pointer<decltype(field)>::type;
    T *tmp = (T *)e->d_stage;
    HIPCHK(sphk::launch_gather(tmp, (const T *)field, 1, 0, idx, ni, e->st));
    HIPCHK(hipMemcpyAsync(field, tmp, sizeof(T) * (size_t)ni,
                          hipMemcpyDeviceToDevice, e->st));
    e->launches += 2;
    return 0;
  };
  for (double *f : {d.px, d.py, d.pz})
    if ((rc = carry(f))) return rc;
  for (float *f : {d.vx, d.vy, d.vz, d.vfx, d.vfy, d.vfz, d.m, d.u, d.ufull,
                   d.h, d.rho, d.rho_dh, d.wc, d.wc_dh, d.divv, d.rotx,
                   d.roty, d.rotz, d.P, d.c, d.f, d.bal, d.vsig, d.lapu,
                   d.avmax, d.alpha, d.adiff, d.divv_prev, d.divv_dt, d.ax,
                   d.ay, d.az, d.udt, d.hdt})
    if ((rc = carry(f))) return rc;
  for (int *f : {d.perm, d.nd, d.ng, d.nf})
    if ((rc = carry(f))) return rc;
  for (int8_t *f : {d.tb, d.mintb})
    if ((rc = carry(f))) return rc;
  HIPCHK(sphk::launch_fill_f32(d.r2b, -1.f, ni, e->st));
  HIPCHK(hipMemsetAsync(e->d_dmax, 0, sizeof(unsigned int), e->st));
  e->dmax = 0.f;
  e->saved_valid = false;
  e->tiles_c = 0;
  e->gen++;
  return finish_binning(e);
}

// ---- time integration ------------------------------------------------------
// get_max_active_bin (src/timeline.h)
int max_active_bin_at(int64_t ti) {
  if (ti == 0) return 56;  // num_time_bins
  int bin = 1;
  while (!((1LL << (bin + 1)) & ti)) ++bin;
  return bin;
}

void set_active_bin(swsph_engine *e, int b) {
  if (e->props.max_active_bin != b) {
    e->props.max_active_bin = b;
    e->gen++;  // captured graphs hold the old activity
  }
}

// Particle range of cell ci, or of all particles for ci == -1.
int part_range(swsph_engine *e, int ci, int *lo, int *hi, int *c_lo,
               int *c_hi) {
  if (ci == -1) {
    *lo = 0;
    *hi = (int)e->n;
    *c_lo = 0;
    *c_hi = e->ncells;
    return 0;
  }
  int rc = cell_ok(e, ci);
  if (rc) return rc;
  *lo = e->cell_start[ci];
  *hi = e->cell_start[ci + 1];
  *c_lo = ci;
  *c_hi = ci + 1;
  return 0;
}

// runner_do_drift_part: drift, then the neighbour records, candidate
// positions and chunk bounds of the drifted particles.
int drift_cells(swsph_engine *e, int ci, int64_t ti_old, int64_t ti_current) {
  int lo, hi, c_lo, c_hi, rc;
  if ((rc = part_range(e, ci, &lo, &hi, &c_lo, &c_hi))) return rc;
  if (ti_current < ti_old) return fail(1, "drift backwards in time");
  DriftArgs a;
  a.dt_drift = (double)(ti_current - ti_old) * e->props.time_base;
  a.dt_kick_hydro = a.dt_drift;
  a.dt_therm = a.dt_drift;
  a.min_u = e->integ.min_u;
  a.g = e->grid;
  a.lo = lo;
  a.hi = hi;
  HIPCHK(sphk::launch_drift(e->d, a, e->d_dmax, e->st));
  e->pos_gen++;
  // every field a record holds may have changed
  for (LoopKind k : {LoopKind::Density, LoopKind::Gradient, LoopKind::Force})
    HIPCHK(sphk::launch_build_records(k, e->d, lo, hi, e->st));
  e->launches += 4;
  e->fields_epoch++;
  if (ci == -1) {
    e->recs_ok = true;
    e->stale_mab = swsph_engine::kNoMab;
  }
  return refresh_bounds(e, c_lo, c_hi);
}

int kick_cells(swsph_engine *e, int ci, int64_t ti_current, int mode) {
  int lo, hi, c_lo, c_hi, rc;
  if ((rc = part_range(e, ci, &lo, &hi, &c_lo, &c_hi))) return rc;
  KickArgs a;
  a.ti_current = ti_current;
  a.time_base = e->props.time_base;
  a.min_u = e->integ.min_u;
  a.mode = mode;
  a.max_active_bin = max_active_bin_at(ti_current);
  a.lo = lo;
  a.hi = hi;
  HIPCHK(sphk::launch_kick(e->d, a, e->st));
  e->launches++;
  if (mode == 2) {
    // hydro_reset_predicted_values changed v, u, P and c
    invalidate_records(e);
    e->fields_epoch++;
  }
  return 0;
}

// runner_do_timestep: new bins of the active particles; returns the next
// synchronisation point over the range and the number of active particles.
int timestep_cells(swsph_engine *e, int ci, int64_t ti_current,
                   int64_t *ti_end_min, int64_t *n_active) {
  int lo, hi, c_lo, c_hi, rc;
  if ((rc = part_range(e, ci, &lo, &hi, &c_lo, &c_hi))) return rc;
  StepArgs a;
  a.ti_current = ti_current;
  a.time_base_inv = 1.0 / e->props.time_base;
  a.dt_min = e->integ.dt_min;
  a.dt_max = e->integ.dt_max;
  a.cfl = e->integ.cfl;
  a.log_max_h_change = e->integ.log_max_h_change;
  a.max_active_bin = max_active_bin_at(ti_current);
  a.fixed = e->integ.fixed_bins;
  a.lo = lo;
  a.hi = hi;
  HIPCHK(sphk::launch_zero_u32(e->d_err, 2, e->st));
  // the minimum starts from max_nr_timesteps
  const unsigned long long top = 1ull << 57;
  HIPCHK(hipMemcpyAsync(e->d_tiend, &top, sizeof(top), hipMemcpyHostToDevice,
                        e->st));
  HIPCHK(sphk::launch_timestep(e->d, a, e->d_tiend, e->d_err, e->st));
  if (!a.fixed) invalidate_records(e);  // force records hold the time bins
  HIPCHK(hipMemcpyAsync(e->h_tiend, e->d_tiend, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipMemcpyAsync(e->h_err, e->d_err, 2 * sizeof(unsigned int),
                        hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  e->launches += 3;
  if (e->h_err[0])
    return fail(3, "%u particles want a time-step below dt_min or of size 0",
                e->h_err[0]);
  if (ti_end_min) *ti_end_min = (int64_t)*e->h_tiend;
  if (n_active) *n_active = (int64_t)e->h_err[1];
  return 0;
}

}  // namespace

// ========================================================================= //
namespace {
// ---- phases of the batched pass ------------------------------------------
// Launches only (no host synchronisation), so that a graph can capture them.

// hydro_init_part on the active particles and the first density round (with
// lists: a fresh neighbour list for every active particle).
int phase_start(swsph_engine *e) {
  HIPCHK(sphk::launch_init_density(e->d, 0, (int)e->n, e->props.max_active_bin,
                                   e->props.h_max, e->st, nullptr, nullptr,
                                   true));
  HIPCHK(sphk::launch_zero_u64(e->d_redo + 3, kGhostWorkSlots, e->st));
  e->launches += 2;
  e->tiles_c = 0;  // activity may have changed since the last pass
  if (e->schedule >= 1)
    return queue_loop(e, LoopKind::Density, 0, nullptr, e->use_lists);
  if (e->use_lists) return build_lists(e, false);
  return stencil(e, LoopKind::Density, false);
}

// The work queue's compact list of the particles its last tile build
// selected (cell order), and its length on the device; null when the pass
// does not run through the queue.  After a density round it holds the
// particles that round processed -- those the ghost finishes -- and after
// the gradient and force loops all active particles.
const int *queue_list(const swsph_engine *e, const int **cnt) {
  if (!e->sweep_list || e->schedule < 1 || e->use_lists || e->n == 0) {
    *cnt = nullptr;
    return nullptr;
  }
  *cnt = e->ts.scan + e->n;
  return e->ts.alist;
}

// One ghost round over the particles of the last density round
// (runner_do_ghost).
int phase_ghost(swsph_engine *e) {
  const int *cnt;
  const int *list = queue_list(e, &cnt);
  GhostArgs ga = ghost_args(e);
  ga.fresh_grad = 1;  // the batched gradient loop starts its own sums
  ga.grad_rec = 1;    // and reads the records the ghost writes
  HIPCHK(sphk::launch_ghost(e->d, ga, 0, (int)e->n, e->d_redo, e->st, list,
                            cnt));
  e->launches += 2;
  return 0;
}

// The density re-run of the particles the last ghost round flagged.  With
// lists it needs the host's counts of that round (redo, relist).
int phase_rerun(swsph_engine *e, int redo, int relist) {
  int rc = 0;
  const bool queue = e->schedule >= 1;
  if (e->use_lists) {
    // Re-runs walk the lists of the flagged particles; the few whose
    // radius outgrew their list go through the chunk loop again.
    if ((rc = list_loop(e, LoopKind::Density, true, redo))) return rc;
    if (relist > 0) {
      e->list_rebuilt += relist;
      return queue ? queue_loop(e, LoopKind::Density, 2, e->d.redo, true)
                   : build_lists(e, true);
    }
    return 0;
  }
  return queue ? queue_loop(e, LoopKind::Density, 1, e->d.redo, false)
               : stencil(e, LoopKind::Density, true);
}

// Gradient records and bounds, gradient loop, extra ghost, force records,
// force loop, end force.
int phase_finish(swsph_engine *e) {
  int rc = 0;
  const int n = (int)e->n;
  const bool queue = e->schedule >= 1;
  // the records of the inactive particles are current (ensure_records),
  // the active ones' gradient records were written by the ghost as each
  // particle converged
  if ((rc = refresh_bounds(e, 0, e->ncells))) return rc;
  if (e->use_lists) {
    if ((rc = list_loop(e, LoopKind::Gradient, false, 0))) return rc;
    if (e->list_fallback > 0 &&
        (rc = queue ? queue_loop(e, LoopKind::Gradient, 1, e->d.lfb, false)
                    : stencil(e, LoopKind::Gradient, true, e->d.lfb)))
      return rc;
  } else if ((rc = queue ? queue_loop(e, LoopKind::Gradient, 0, nullptr, false)
                         : stencil(e, LoopKind::Gradient, false))) {
    return rc;
  }
  // after the gradient loop the queue's list holds all active particles
  const int *cnt;
  const int *list = queue_list(e, &cnt);
  GhostArgs gx = ghost_args(e);
  gx.fresh_force = 1;
  // With the narrow tiles of sparse activity the extra ghost writes the
  // active force records as it goes; otherwise (most particles active) the
  // builder's wave-staged contiguous stores are faster than one scattered
  // 96-byte record per thread.
  const bool narrow = e->schedule >= 1 &&
                      e->loop_c[(int)LoopKind::Force] == kMinTile;
  gx.force_rec = narrow ? 1 : 0;
  HIPCHK(sphk::launch_extra_ghost(e->d, gx, 0, n, e->st, list, cnt));
  e->launches++;
  if (!narrow) {
    HIPCHK(sphk::launch_build_records(LoopKind::Force, e->d, 0, n, e->st,
                                      e->props.max_active_bin, list, cnt));
    e->launches++;
  }
  if ((rc = queue ? queue_loop(e, LoopKind::Force, 0, nullptr, false)
                  : stencil(e, LoopKind::Force, false)))
    return rc;
  HIPCHK(sphk::launch_end_force(e->d, 0, n, e->props.max_active_bin, e->st,
                                list, cnt));
  e->launches++;
  return 0;
}

int finish_stats(swsph_engine *e, swsph_pass_stats *stats, int rounds,
                 int64_t l0) {
  int rc = 0;
  if (stats) {
    stats->ghost_rounds = rounds;
    stats->launches = e->launches - l0;
    stats->list_rebuilt = e->list_rebuilt;
    stats->list_fallback = e->list_fallback;
    stats->n_density_all = e->density_work;
  }
  if ((rc = count_into(e, stats))) return rc;
  if ((rc = check_support(e))) return rc;
  return collect_timing(e);
}

// Captures the launches of fn on the engine stream into an executable graph.
template <class F>
int capture(swsph_engine *e, hipGraphExec_t *out, F fn) {
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
  e->capturing = true;
  const int rc = fn();
  e->capturing = false;
  const hipError_t ec = hipStreamEndCapture(e->st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return fail(2, "graph capture: %s", hipGetErrorString(ec));
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess)
    return fail(2, "graph instantiation: %s", hipGetErrorString(ei));
  return 0;
}

void drop_graphs(swsph_engine *e) {
  if (e->g_density) (void)hipGraphExecDestroy(e->g_density);
  if (e->g_finish) (void)hipGraphExecDestroy(e->g_finish);
  e->g_density = e->g_finish = nullptr;
}

// A set of nf device buffers holding at least n particles each; cap is the
// particle count the set was allocated for (reallocated when it is short).
int ensure_copies(swsph_engine *e, std::vector<float *> &buf, size_t &cap,
                  size_t nf, size_t n) {
  if (!buf.empty() && cap >= n && buf.size() == nf) return 0;
  for (float *p : buf)
    if (p) (void)hipFree(p);
  buf.assign(nf, nullptr);
  cap = 0;
  const size_t want = std::max(n, e->cap_n);
  for (auto &p : buf) HIPCHK(dalloc(&p, want));
  cap = want;
  return 0;
}

// Device copies of the pass state (h and the (io) fields) into / out of
// the autotuner's buffers.
int copy_state(swsph_engine *e, std::vector<float *> &buf, bool save) {
  std::vector<float *> f = io_fields(e->d);
  const size_t n = (size_t)(e->n ? e->n : 1);
  if (save) {
    int rc = ensure_copies(e, buf, e->tune_cap, f.size(), n);
    if (rc) return rc;
  } else if (buf.size() != f.size() || e->tune_cap < n) {
    return fail(2, "autotuner state copy does not hold the current particles");
  }
  // all fields in one launch
  sphk::CopySet cs{};
  cs.m = (int)f.size();
  if (cs.m > sphk::kMaxCopies) return fail(2, "too many state fields");
  for (int k = 0; k < cs.m; k++) {
    cs.src[k] = save ? f[k] : buf[k];
    cs.dst[k] = save ? buf[k] : f[k];
  }
  HIPCHK(sphk::launch_copy_set(cs, (int)n, e->st));
  if (!save) e->stale_mab = e->last_pass_mab;
  return 0;
}

// A pass over no particles: nothing to do, zero counts.
int empty_pass(swsph_pass_stats *stats) {
  if (stats) *stats = swsph_pass_stats{};
  return 0;
}

// The eager batched pass (the engine mutex held by the caller).
int pass_batched_locked(swsph_engine *e, swsph_pass_stats *stats) {
  int rc = 0;
  if (e->need_particles)
    return fail(1, "the grid changed: call swsph_set_particles first");
  if (e->n == 0) return empty_pass(stats);
  const int64_t l0 = e->launches;
  e->list_rebuilt = 0;
  e->list_fallback = 0;
  e->density_work = 0;
  if (e->use_lists && (rc = ensure_lists(e))) return rc;
  if ((rc = ensure_records(e))) return rc;
  e->last_pass_mab = e->props.max_active_bin;
  if ((rc = phase_start(e))) return rc;
  int redo = 0, it = 0;
  for (; it < e->props.max_iter; it++) {
    if ((rc = phase_ghost(e))) return rc;
    int relist = 0, nolist = 0;
    if ((rc = read_redo(e, &redo, &relist, &nolist))) return rc;
    e->list_fallback += nolist;
    if (redo == 0) break;
    if ((rc = phase_rerun(e, redo, relist))) return rc;
  }
  if (redo > 0)
    return fail(3, "Smoothing length failed to converge on %d particles", redo);
  e->last_rounds = it + 1;
  if ((rc = phase_finish(e))) return rc;
  return finish_stats(e, stats, it + 1, l0);
}
// Device arrays of a pass output or state field: up to three components
// of one element type, in slot order.
struct FieldView {
  enum Kind { F64, F32, I32, I8 } kind = F32;
  const void *c[3] = {nullptr, nullptr, nullptr};
  int ncomp = 1;
  int esize() const { return kind == F64 ? 8 : (kind == I8 ? 1 : 4); }
};

bool field_view(const DevParts &d, int field, FieldView &v) {
  auto one = [&](FieldView::Kind k, const void *p) {
    v.kind = k;
    v.c[0] = p;
    v.ncomp = 1;
    return true;
  };
  auto three = [&](FieldView::Kind k, const void *a, const void *b,
                   const void *c) {
    v.kind = k;
    v.c[0] = a;
    v.c[1] = b;
    v.c[2] = c;
    v.ncomp = 3;
    return true;
  };
  const auto F = FieldView::F32;
  switch (field) {
    case SWSPH_H: return one(F, d.h);
    case SWSPH_RHO: return one(F, d.rho);
    case SWSPH_RHO_DH: return one(F, d.rho_dh);
    case SWSPH_WCOUNT: return one(F, d.wc);
    case SWSPH_WCOUNT_DH: return one(F, d.wc_dh);
    case SWSPH_DIV_V: return one(F, d.divv);
    case SWSPH_ROT_V: return three(F, d.rotx, d.roty, d.rotz);
    case SWSPH_PRESSURE: return one(F, d.P);
    case SWSPH_SOUNDSPEED: return one(F, d.c);
    case SWSPH_F: return one(F, d.f);
    case SWSPH_BALSARA: return one(F, d.bal);
    case SWSPH_V_SIG: return one(F, d.vsig);
    case SWSPH_LAPLACE_U: return one(F, d.lapu);
    case SWSPH_ALPHA_VISC_MAX_NGB: return one(F, d.avmax);
    case SWSPH_ALPHA_VISC: return one(F, d.alpha);
    case SWSPH_ALPHA_DIFF: return one(F, d.adiff);
    case SWSPH_DIV_V_DT: return one(F, d.divv_dt);
    case SWSPH_DIV_V_PREV: return one(F, d.divv_prev);
    case SWSPH_A_HYDRO: return three(F, d.ax, d.ay, d.az);
    case SWSPH_U_DT: return one(F, d.udt);
    case SWSPH_H_DT: return one(F, d.hdt);
    case SWSPH_N_DENSITY: return one(FieldView::I32, d.nd);
    case SWSPH_N_GRADIENT: return one(FieldView::I32, d.ng);
    case SWSPH_N_FORCE: return one(FieldView::I32, d.nf);
    case SWSPH_MIN_NGB_TIME_BIN: return one(FieldView::I8, d.mintb);
    case SWSPH_TIME_BIN: return one(FieldView::I8, d.tb);
    case SWSPH_X: return three(FieldView::F64, d.px, d.py, d.pz);
    case SWSPH_V: return three(F, d.vx, d.vy, d.vz);
    case SWSPH_V_FULL: return three(F, d.vfx, d.vfy, d.vfz);
    case SWSPH_U: return one(F, d.u);
    case SWSPH_U_FULL: return one(F, d.ufull);
    default: return false;
  }
}
// ---- bundles of cell tasks (ghost, extra ghost, end force, init) --------
// The cells of a bundle as the device mask that the work-queue selections
// read, until clear_cells.
int select_cells(swsph_engine *e, const int *cells, int n) {
  if (n < 0 || (n > 0 && !cells)) return fail(1, "bad cell list");
  if (n > e->ncells) return fail(1, "%d cells listed, %d exist", n, e->ncells);
  for (int k = 0; k < n; k++)
    if (int rc = cell_ok(e, cells[k])) return rc;
  if (!e->d_cmask || e->cmask_cap < e->ncells) {
    for (void *p : {(void *)e->d_cmask, (void *)e->d_cact, (void *)e->d_clist})
      if (p) (void)hipFree(p);
    e->d_cmask = e->d_cact = nullptr;
    e->d_clist = nullptr;
    e->cmask_cap = 0;
    HIPCHK(dalloc(&e->d_cmask, (size_t)e->ncells));
    HIPCHK(dalloc(&e->d_cact, (size_t)e->ncells));
    HIPCHK(dalloc(&e->d_clist, (size_t)e->ncells));
    e->cmask_cap = e->ncells;
  }
  if (n > 0)
    HIPCHK(hipMemcpyAsync(e->d_clist, cells, sizeof(int) * (size_t)n,
                          hipMemcpyHostToDevice, e->st));
  HIPCHK(sphk::launch_cell_mask(e->d_cmask, e->ncells, e->d_clist, n, e->st));
  e->launches += 2;
  // (the host list must outlive its copy)
  HIPCHK(hipStreamSynchronize(e->st));
  e->cmask_on = true;
  e->tiles_c = 0;
  return 0;
}

void clear_cells(swsph_engine *e) {
  e->cmask_on = false;
  e->tiles_c = 0;
}

// The active particles of the selected cells (mode 0), or those of them
// flagged in `flag` (mode 1), as the work queue's compact list and its
// device-side length.
int select_particles(swsph_engine *e, int mode, const int *flag,
                     const int **list, const int **cnt) {
  HIPCHK(sphk::launch_build_tiles(e->d, loop_args(e), (int)e->n, e->ncells,
                                  16, mode, flag, e->ts, e->st));
  e->launches += 12;
  e->tiles_c = 0;
  *list = e->ts.alist;
  *cnt = e->ts.scan + e->n;
  return 0;
}

// hydro_init_part on the active particles of the cells.
int init_cells_locked(swsph_engine *e, const int *cells, int n) {
  int rc;
  if ((rc = ensure_records(e)) || (rc = select_cells(e, cells, n))) return rc;
  const int *list, *cnt;
  if (!(rc = select_particles(e, 0, nullptr, &list, &cnt))) {
    const hipError_t err = sphk::launch_init_density(
        e->d, 0, (int)e->n, e->props.max_active_bin, e->props.h_max, e->st,
        list, cnt);
    e->launches++;
    if (err != hipSuccess) rc = fail(2, "init: %s", hipGetErrorString(err));
  }
  clear_cells(e);
  return rc;
}

// runner_do_ghost on each of the cells, as device-wide rounds over their
// particles: end of density, Newton-Raphson / bisection step, and the
// density re-run of the particles that go again over their 27-cell
// stencil; then their gradient records and the chunk bounds.
int ghost_cells_locked(swsph_engine *e, const int *cells, int n, int *rounds) {
  int rc;
  if ((rc = select_cells(e, cells, n))) return rc;
  const GhostArgs ga = ghost_args(e);
  const int *list, *cnt;
  int redo = 0, it = 0;
  for (; it < e->props.max_iter; it++) {
    if ((rc = select_particles(e, 1, e->d.redo, &list, &cnt))) break;
    const hipError_t err = sphk::launch_ghost(e->d, ga, 0, (int)e->n,
                                              e->d_redo, e->st, list, cnt);
    e->launches += 2;
    if (err != hipSuccess) {
      rc = fail(2, "ghost: %s", hipGetErrorString(err));
      break;
    }
    if ((rc = read_redo(e, &redo)) || redo == 0) break;
    if ((rc = queue_loop(e, LoopKind::Density, 1, e->d.redo, false))) break;
  }
  if (!rc && redo > 0)
    rc = fail(3, "Smoothing length failed to converge on %d particles", redo);
  if (!rc && !(rc = select_particles(e, 0, nullptr, &list, &cnt))) {
    const hipError_t err = sphk::launch_build_records(
        LoopKind::Gradient, e->d, 0, (int)e->n, e->st,
        e->props.max_active_bin, list, cnt);
    e->launches++;
    if (err != hipSuccess) rc = fail(2, "records: %s", hipGetErrorString(err));
  }
  clear_cells(e);
  if (!rc) rc = refresh_bounds(e, 0, e->ncells);
  if (rounds) *rounds = std::max(*rounds, it + 1);
  return rc;
}

// runner_do_extra_ghost (then the force records) or runner_do_end_hydro_-
// force on the active particles of the cells.
int cell_op_locked(swsph_engine *e, const int *cells, int n, bool extra) {
  int rc;
  if ((rc = select_cells(e, cells, n))) return rc;
  const int *list, *cnt;
  if (!(rc = select_particles(e, 0, nullptr, &list, &cnt))) {
    hipError_t err;
    if (extra) {
      err = sphk::launch_extra_ghost(e->d, ghost_args(e), 0, (int)e->n,
                                     e->st, list, cnt);
      if (err == hipSuccess)
        err = sphk::launch_build_records(LoopKind::Force, e->d, 0, (int)e->n,
                                         e->st, e->props.max_active_bin, list,
                                         cnt);
      e->launches += 2;
    } else {
      err = sphk::launch_end_force(e->d, 0, (int)e->n,
                                   e->props.max_active_bin, e->st, list, cnt);
      e->launches++;
    }
    if (err != hipSuccess) rc = fail(2, "cell tasks: %s", hipGetErrorString(err));
  }
  clear_cells(e);
  return rc;
}

// A bundle of released self / pair tasks of one loop (swsph_run_tasks).
int run_tasks_locked(swsph_engine *e, int loop, const swsph_task *tasks,
                     int ntasks) {
  if (!tasks && ntasks > 0) return fail(1, "null argument");
  if (loop < 0 || loop > 2) return fail(1, "loop index %d out of range", loop);
  if (ntasks < 0) return fail(1, "negative task count");
  if (ntasks == 0) return 0;
  if (e->need_particles) return fail(1, "no particles for the current grid");
  const LoopKind k = (LoopKind)loop;
  // Per i-cell, bit o of the stencil images it meets (o = 9 (ox + 1) +
  // 3 (oy + 1) + (oz + 1) for the offset of the image's cell), and the
  // i-cells in the order the tasks first touch them.
  std::vector<uint32_t> mask((size_t)e->ncells, 0u);
  std::vector<int> cells;
  auto add = [&](int ci, int o, int cj) -> int {
    if (!mask[ci]) cells.push_back(ci);
    if (mask[ci] >> o & 1u)
      return fail(1, "cell %d meets cell %d twice in one bundle", ci, cj);
    mask[ci] |= 1u << o;
    return 0;
  };
  const int *nc = e->grid.nc;
  auto coords = [&](int c, int v[3]) {
    v[0] = c / (nc[1] * nc[2]);
    v[1] = (c / nc[2]) % nc[1];
    v[2] = c % nc[2];
  };
  int rc = 0;
  for (int n = 0; n < ntasks; n++) {
    const swsph_task &tk = tasks[n];
    if ((rc = cell_ok(e, tk.ci))) return rc;
    if (tk.type == SWSPH_TASK_SELF) {
      if ((rc = add(tk.ci, 13, tk.ci))) return rc;
    } else if (tk.type == SWSPH_TASK_PAIR) {
      if ((rc = cell_ok(e, tk.cj))) return rc;
      double t[3];
      if (tk.ci == tk.cj || !adjacent(e, tk.ci, tk.cj, t))
        return fail(1, "cells %d and %d are not neighbours", tk.ci, tk.cj);
      int a[3], b[3], o = 0;
      coords(tk.ci, a);
      coords(tk.cj, b);
      for (int q = 0; q < 3; q++) {
        int dd = b[q] - a[q];
        dd = dd > 1 ? dd - nc[q] : (dd < -1 ? dd + nc[q] : dd);
        o = 3 * o + dd + 1;
      }
      if ((rc = add(tk.ci, o, tk.cj)) || (rc = add(tk.cj, 26 - o, tk.ci)))
        return rc;
    } else {
      return fail(1, "task %d: unknown type %d", n, tk.type);
    }
  }
  if ((rc = ensure_records(e))) return rc;
  const swsph_engine::ChunkTable *tb = chunk_table(e, chunk_c(e, loop), &rc);
  if (!tb) return rc;
  std::vector<BundleItem> items;
  std::vector<BundleNbr> nbrs;
  for (const int c : cells) {
    // all 27 images: the whole stencil, no list
    const bool whole = mask[c] == (1u << 27) - 1u;
    const int off = (int)nbrs.size();
    int cnt = 0;
    if (!whole) {
      int a[3];
      coords(c, a);
      for (int o = 0; o < 27; o++) {
        if (!(mask[c] >> o & 1u)) continue;
        const int d[3] = {o / 9 - 1, (o / 3) % 3 - 1, o % 3 - 1};
        BundleNbr bn{};
        int v[3];
        for (int q = 0; q < 3; q++) {
          v[q] = a[q] + d[q];
          bn.t[q] = 0.0;
          if (v[q] < 0) {
            v[q] += nc[q];
            bn.t[q] = -e->grid.dim[q];
          } else if (v[q] >= nc[q]) {
            v[q] -= nc[q];
            bn.t[q] = e->grid.dim[q];
          }
        }
        bn.cell = (v[0] * nc[1] + v[1]) * nc[2] + v[2];
        nbrs.push_back(bn);
        cnt++;
      }
    }
    for (int ch = tb->cell_chunk_h[c]; ch < tb->cell_chunk_h[c + 1]; ch++)
      items.push_back(BundleItem{ch, off, whole ? -1 : cnt, 0});
  }
  if (items.empty()) return 0;
  if (items.size() > e->cap_bitems) {
    if (e->d_bitems) (void)hipFree(e->d_bitems);
    e->d_bitems = nullptr;
    HIPCHK(hipMalloc((void **)&e->d_bitems, items.size() * sizeof(BundleItem)));
    e->cap_bitems = items.size();
  }
  if (nbrs.empty()) nbrs.push_back(BundleNbr{});
  if (nbrs.size() > e->cap_bnbrs) {
    if (e->d_bnbrs) (void)hipFree(e->d_bnbrs);
    e->d_bnbrs = nullptr;
    HIPCHK(hipMalloc((void **)&e->d_bnbrs, nbrs.size() * sizeof(BundleNbr)));
    e->cap_bnbrs = nbrs.size();
  }
  // the host vectors must outlive the copies: wait for them
  HIPCHK(hipMemcpyAsync(e->d_bitems, items.data(),
                        items.size() * sizeof(BundleItem),
                        hipMemcpyHostToDevice, e->st));
  HIPCHK(hipMemcpyAsync(e->d_bnbrs, nbrs.data(), nbrs.size() * sizeof(BundleNbr),
                        hipMemcpyHostToDevice, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  BundleArgs ba{e->d_bitems, e->d_bnbrs, (int)items.size()};
  HIPCHK(sphk::launch_bundle(k, e->d, loop_args(e), chunk_args(tb), ba, e->st));
  e->launches++;
  return 0;
}


// The engine as the device side of the task scheduler (sph_sched.cpp); the
// engine mutex is held by the caller.
struct EngineTasks final : sphs::DeviceTasks {
  swsph_engine *e;
  explicit EngineTasks(swsph_engine *eng) : e(eng) {}
  int cell_active(std::vector<unsigned char> &act) override {
    int rc = select_cells(e, nullptr, 0);  // allocates the masks
    clear_cells(e);
    if (rc) return rc;
    HIPCHK(sphk::launch_cell_active(e->d, (int)e->n, e->ncells,
                                    e->props.max_active_bin, e->d_cact,
                                    e->st));
    e->launches += 2;
    act.assign((size_t)e->ncells, 0);
    HIPCHK(hipMemcpyAsync(act.data(), e->d_cact, (size_t)e->ncells,
                          hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    return 0;
  }
  int init_cells(const std::vector<int> &c) override {
    return init_cells_locked(e, c.data(), (int)c.size());
  }
  int loop_tasks(int loop, const std::vector<swsph_task> &t) override {
    return run_tasks_locked(e, loop, t.data(), (int)t.size());
  }
  int ghost_cells(const std::vector<int> &c, int *rounds) override {
    return ghost_cells_locked(e, c.data(), (int)c.size(), rounds);
  }
  int extra_ghost_cells(const std::vector<int> &c) override {
    return cell_op_locked(e, c.data(), (int)c.size(), true);
  }
  int end_force_cells(const std::vector<int> &c) override {
    return cell_op_locked(e, c.data(), (int)c.size(), false);
  }
};

// One pass through the task graph (swsph_sched_pass).
int sched_pass_locked(swsph_engine *e, swsph_sched_stats *ss,
                      swsph_pass_stats *stats) {
  int rc;
  if (e->need_particles)
    return fail(1, "the grid changed: call swsph_set_particles first");
  if (e->n == 0) {
    if (ss) *ss = swsph_sched_stats{};
    return empty_pass(stats);
  }
  const int64_t l0 = e->launches;
  e->list_rebuilt = e->list_fallback = 0;
  e->density_work = 0;
  if ((rc = ensure_records(e))) return rc;
  e->last_pass_mab = e->props.max_active_bin;
  HIPCHK(sphk::launch_zero_u64(e->d_redo + 3, kGhostWorkSlots, e->st));
  e->launches++;
  EngineTasks dev(e);
  int rounds = 0;
  swsph_sched_stats local{};
  rc = sphs::run_hydro_graph(dev, e->grid.nc, ss ? ss : &local, &rounds);
  clear_cells(e);
  if (rc) return rc == 2 ? fail(2, "task graph left tasks unreleased") : rc;
  e->last_rounds = std::max(1, rounds);
  return finish_stats(e, stats, std::max(1, rounds), l0);
}
}  // namespace

extern "C" {

const char *swsph_last_error(void) { return g_err.c_str(); }

int swsph_engine_create(const swsph_props *props, swsph_engine **out) {
  if (!props || !out) return fail(1, "null argument");
  int rc = check_props(props);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(5, "no HIP device available");
  swsph_engine *e = new swsph_engine();
  e->props = *props;
  for (int k = 0; k < 3; k++) {
    e->grid.nc[k] = props->ncell[k];
    e->grid.dim[k] = props->dim[k];
  }
  e->ncells = props->ncell[0] * props->ncell[1] * props->ncell[2];
  if (hipGetDevice(&e->device) != hipSuccess) {
    delete e;
    return fail(2, "hipGetDevice failed");
  }
  // the fixed scratch of the allocation-free pass on this device
  if (hipError_t err = sphk::noalloc_setup(e->device)) {
    delete e;
    return fail(2, "device address of the scratch array: %s",
                hipGetErrorString(err));
  }
  if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return fail(2, "hipStreamCreate failed");
  }
  if (hipMalloc((void **)&e->d_redo,
                kGhostCounters * sizeof(unsigned long long)) !=
          hipSuccess ||
      hipMalloc((void **)&e->d_hist, swsph_engine::kHistRounds *
                                         sizeof(unsigned long long)) !=
          hipSuccess ||
      hipHostMalloc((void **)&e->h_hist, swsph_engine::kHistRounds *
                                             sizeof(unsigned long long),
                    0) != hipSuccess ||
      hipHostMalloc((void **)&e->h_redo,
                    kGhostCounters * sizeof(unsigned long long), 0) !=
          hipSuccess ||
      hipMalloc((void **)&e->d_cnt, 3 * sizeof(unsigned long long)) != hipSuccess ||
      hipHostMalloc((void **)&e->h_cnt, 3 * sizeof(unsigned long long), 0) !=
          hipSuccess ||
      hipMalloc((void **)&e->d_hg2_all, sizeof(unsigned int)) != hipSuccess ||
      hipHostMalloc((void **)&e->h_hg2_all, sizeof(unsigned int), 0) !=
          hipSuccess ||
      hipMalloc((void **)&e->d_dmax, sizeof(unsigned int)) != hipSuccess ||
      hipMalloc((void **)&e->d_tiend, sizeof(unsigned long long)) !=
          hipSuccess ||
      hipMalloc((void **)&e->d_err, 3 * sizeof(unsigned int)) != hipSuccess ||
      hipHostMalloc((void **)&e->h_small, 2 * sizeof(unsigned int), 0) !=
          hipSuccess ||
      hipHostMalloc((void **)&e->h_tiend, sizeof(unsigned long long), 0) !=
          hipSuccess ||
      hipHostMalloc((void **)&e->h_err, 3 * sizeof(unsigned int), 0) !=
          hipSuccess ||
      hipHostMalloc((void **)&e->h_ntiles, sizeof(int), 0) != hipSuccess) {
    delete e;
    return fail(2, "allocation of engine counters failed");
  }
  // SPH:max_volume_change default 1.4 (src/hydro_properties.c:42, 122)
  e->integ.log_max_h_change = logf(powf(1.4f, 0.3333333333f));
  // One table per admissible chunk capacity at most: never reallocated, so
  // table pointers stay valid.
  e->tables.reserve(4);
  *out = e;
  return 0;
}

int swsph_engine_destroy(swsph_engine *e) {
  if (!e) return 0;
  (void)hipStreamSynchronize(e->st);
  free_dev(e);
  for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
  if (e->g_density) (void)hipGraphExecDestroy(e->g_density);
  if (e->g_finish) (void)hipGraphExecDestroy(e->g_finish);
  if (e->d_redo) (void)hipFree(e->d_redo);
  if (e->h_redo) (void)hipHostFree(e->h_redo);
  if (e->d_hist) (void)hipFree(e->d_hist);
  if (e->h_hist) (void)hipHostFree(e->h_hist);
  if (e->d_cnt) (void)hipFree(e->d_cnt);
  if (e->d_hg2_all) (void)hipFree(e->d_hg2_all);
  if (e->h_hg2_all) (void)hipHostFree(e->h_hg2_all);
  if (e->h_cnt) (void)hipHostFree(e->h_cnt);
  if (e->d_dmax) (void)hipFree(e->d_dmax);
  if (e->d_tiend) (void)hipFree(e->d_tiend);
  if (e->d_err) (void)hipFree(e->d_err);
  if (e->d_bitems) (void)hipFree(e->d_bitems);
  if (e->d_bnbrs) (void)hipFree(e->d_bnbrs);
  if (e->h_small) (void)hipHostFree(e->h_small);
  if (e->h_tiend) (void)hipHostFree(e->h_tiend);
  if (e->h_err) (void)hipHostFree(e->h_err);
  if (e->h_ntiles) (void)hipHostFree(e->h_ntiles);
  if (e->st) (void)hipStreamDestroy(e->st);
  delete e;
  return 0;
}

int swsph_set_props(swsph_engine *e, const swsph_props *props) {
  if (!e || !props) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  e->gen++;  // captured graphs no longer match
  int rc = check_props(props);
  if (rc) return rc;
  bool grid = false;
  for (int k = 0; k < 3; k++)
    grid = grid || props->ncell[k] != e->props.ncell[k] ||
           props->dim[k] != e->props.dim[k];
  e->props = *props;
  e->pos_gen++;  // the pre-test records depend on the grid and pos_eps
  if (grid) {
    // the particles must be binned again (swsph_set_particles), which
    // reuses the device buffers when they are large enough
    for (int k = 0; k < 3; k++) {
      e->grid.nc[k] = props->ncell[k];
      e->grid.dim[k] = props->dim[k];
    }
    e->ncells = props->ncell[0] * props->ncell[1] * props->ncell[2];
    e->need_particles = e->n > 0 || !e->cell_start.empty();
    for (auto &t : e->tables) t.valid = false;
  }
  return 0;
}

int swsph_set_particles(swsph_engine *e, const swsph_host_parts *p) {
  if (!e || !p) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  e->gen++;  // captured graphs no longer match
  const int64_t n = p->n;
  // Neighbour-loop row tables pack a particle index in 27 bits.
  if (n < 0 || n > ((int64_t)1 << 27))
    return fail(1, "particle count %lld out of range (max 2^27)", (long long)n);
  if (!p->x || !p->v || !p->m || !p->u || !p->time_bin || !p->h)
    return fail(1, "missing particle arrays");
  // Positions must lie in the box (the reference wraps them when it
  // rebuilds its cells); the FP32 pre-test bound relies on every particle
  // lying in its top-level cell.
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      const double xk = p->x[3 * i + k];
      if (!(xk >= 0.0 && xk < e->grid.dim[k]))
        return fail(1, "particle %lld lies outside the periodic box",
                    (long long)i);
    }
  HIPCHK(hipStreamSynchronize(e->st));
  int rc = 0;
  if (!e->d.px || (size_t)n > e->cap_n || e->ncells > e->cap_cells) {
    free_dev(e);
    if ((rc = alloc_dev(e, (size_t)std::max<int64_t>(n, 1)))) return rc;
  }
  e->n = n;
  e->need_particles = false;
  e->started = false;  // a new particle set starts a new time line
  e->ntiles = 0;
  if ((rc = reset_dev(e, (size_t)n))) return rc;
  const int ni = (int)n;
  DevParts &d = e->d;
  if (n > 0) {
    // Positions in the caller's order to the staging buffer; keys, radix
    // sort and cell bounds on the device; the sort's values are the
    // caller's indices, i.e. the permutation sorted slot -> caller index.
    HIPCHK(hipMemcpyAsync(e->d_stage, p->x, sizeof(double) * 3 * n,
                          hipMemcpyHostToDevice, e->st));
    // the time bins in the caller's order for the keys (d.tb is written in
    // slot order from the staging buffer below)
    if (e->bin_order)
      HIPCHK(hipMemcpyAsync(d.tb, p->time_bin, n, hipMemcpyHostToDevice,
                            e->st));
    if ((rc = bin_sort(e, e->d_stage, e->d_stage + 1, e->d_stage + 2, 3,
                       d.perm, e->bin_order ? d.tb : nullptr)))
      return rc;
    double *pd[3] = {d.px, d.py, d.pz};
    for (int k = 0; k < 3; k++)
      HIPCHK(sphk::launch_gather(pd[k], (const double *)e->d_stage, 3, k,
                                 d.perm, ni, e->st));
    // Velocities (predicted and v_full start equal, as in
    // hydro_first_init_part), then the scalar fields, each through the
    // staging buffer (stream order keeps uploads and gathers apart).
    float *sf = (float *)e->d_stage;
    HIPCHK(hipMemcpyAsync(sf, p->v, sizeof(float) * 3 * n,
                          hipMemcpyHostToDevice, e->st));
    float *pv[3] = {d.vx, d.vy, d.vz}, *pvf[3] = {d.vfx, d.vfy, d.vfz};
    for (int k = 0; k < 3; k++) {
      HIPCHK(sphk::launch_gather(pv[k], (const float *)sf, 3, k, d.perm, ni,
                                 e->st));
      HIPCHK(sphk::launch_gather(pvf[k], (const float *)sf, 3, k, d.perm, ni,
                                 e->st));
    }
    struct Up {
      const float *src;
      float *dst, *dst2;
      float dflt;
    } ups[] = {{p->m, d.m, nullptr, 0.f},
               {p->u, d.u, d.ufull, 0.f},
               {p->h, d.h, nullptr, 0.f},
               {p->rho, d.rho, nullptr, 0.f},
               {p->pressure, d.P, nullptr, 0.f},
               {p->soundspeed, d.c, nullptr, 0.f},
               {p->f, d.f, nullptr, 0.f},
               {p->balsara, d.bal, nullptr, 0.f},
               {p->alpha_visc, d.alpha, nullptr, 0.1f},
               {p->alpha_diff, d.adiff, nullptr, 0.f},
               {p->div_v_prev, d.divv_prev, nullptr, 0.f}};
    for (const Up &u : ups) {
      if (!u.src) {
        HIPCHK(sphk::launch_fill_f32(u.dst, u.dflt, ni, e->st));
        if (u.dst2) HIPCHK(sphk::launch_fill_f32(u.dst2, u.dflt, ni, e->st));
        continue;
      }
      HIPCHK(hipMemcpyAsync(sf, u.src, sizeof(float) * n,
                            hipMemcpyHostToDevice, e->st));
      HIPCHK(sphk::launch_gather(u.dst, (const float *)sf, 1, 0, d.perm, ni,
                                 e->st));
      if (u.dst2)
        HIPCHK(sphk::launch_gather(u.dst2, (const float *)sf, 1, 0, d.perm,
                                   ni, e->st));
    }
    int8_t *sb = (int8_t *)e->d_stage;
    HIPCHK(hipMemcpyAsync(sb, p->time_bin, n, hipMemcpyHostToDevice, e->st));
    HIPCHK(sphk::launch_gather(d.tb, (const int8_t *)sb, 1, 0, d.perm, ni,
                               e->st));
  } else {
    HIPCHK(sphk::launch_cell_bounds(e->d_key[1], 0, e->ncells,
                                    e->d_cell_start, d.cell, e->st, 0));
  }
  if ((rc = finish_binning(e))) return rc;
  return check_fits(e);
}

int swsph_get_field(swsph_engine *e, int field, void *out) {
  if (!e || !out) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  const int n = (int)e->n;
  FieldView v;
  if (!field_view(e->d, field, v)) return fail(1, "unknown field %d", field);
  // Scattered into the caller's order on the device (through the staging
  // buffer, 24 B per particle), then one contiguous copy to the host.
  auto fetch = [&](auto tag) -> int {
    using T = decltype(tag);
    T *stage = (T *)e->d_stage;
    for (int k = 0; k < v.ncomp; k++)
      HIPCHK(sphk::launch_scatter(stage, (const T *)v.c[k], v.ncomp, k,
                                  e->d.perm, n, e->st));
    if (n)
      HIPCHK(hipMemcpyAsync(out, stage, sizeof(T) * (size_t)v.ncomp * n,
                            hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    return 0;
  };
  switch (v.kind) {
    case FieldView::F64: return fetch(double{});
    case FieldView::F32: return fetch(float{});
    case FieldView::I32: return fetch(int{});
    default: return fetch(int8_t{});
  }
}

int swsph_get_fields(swsph_engine *e, int nf, const int *fields,
                     void *const *outs) {
  if (!e || nf < 0 || (nf > 0 && (!fields || !outs)))
    return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  const int n = (int)e->n;
  for (int q = 0; q < nf; q++) {
    FieldView v;
    if (!field_view(e->d, fields[q], v))
      return fail(1, "unknown field %d", fields[q]);
    if (!outs[q]) return fail(1, "null output for field %d", fields[q]);
    // scattered into the caller's order in the staging buffer, copied out;
    // the stream keeps the next field's scatter behind this copy
    auto fetch = [&](auto tag) -> int {
      using T = decltype(tag);
      T *stage = (T *)e->d_stage;
      for (int k = 0; k < v.ncomp; k++)
        HIPCHK(sphk::launch_scatter(stage, (const T *)v.c[k], v.ncomp, k,
                                    e->d.perm, n, e->st));
      if (n)
        HIPCHK(hipMemcpyAsync(outs[q], stage, sizeof(T) * (size_t)v.ncomp * n,
                              hipMemcpyDeviceToHost, e->st));
      return 0;
    };
    int rc;
    switch (v.kind) {
      case FieldView::F64: rc = fetch(double{}); break;
      case FieldView::F32: rc = fetch(float{}); break;
      case FieldView::I32: rc = fetch(int{}); break;
      default: rc = fetch(int8_t{});
    }
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(e->st));
  return 0;
}

int swsph_cell_write_back(swsph_engine *e, int ci, const swsph_field_dst *dst,
                          int nfields) {
  if (!e || (nfields > 0 && !dst) || nfields < 0)
    return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  int lo = 0, hi = (int)e->n;
  if (ci != -1) {
    if (int rc = cell_ok(e, ci)) return rc;
    lo = e->cell_start[ci];
    hi = e->cell_start[ci + 1];
  }
  std::vector<FieldView> views((size_t)nfields);
  for (int f = 0; f < nfields; f++) {
    if (!field_view(e->d, dst[f].field, views[f]))
      return fail(1, "unknown field %d", dst[f].field);
    const int64_t need = (int64_t)views[f].ncomp * views[f].esize();
    if (!dst[f].base || std::llabs((long long)dst[f].stride) < need)
      return fail(1, "destination of field %d: null base or stride %lld "
                  "below its %lld bytes", dst[f].field,
                  (long long)dst[f].stride, (long long)need);
  }
  // completion: everything enqueued so far has finished
  HIPCHK(hipStreamSynchronize(e->st));
  const size_t cnt = (size_t)(hi - lo);
  if (cnt == 0 || nfields == 0) return 0;
  std::vector<int> perm(cnt);
  HIPCHK(hipMemcpy(perm.data(), e->d.perm + lo, sizeof(int) * cnt,
                   hipMemcpyDeviceToHost));
  std::vector<unsigned char> stage(cnt * 8);
  for (int f = 0; f < nfields; f++) {
    const FieldView &v = views[f];
    const size_t es = (size_t)v.esize();
    unsigned char *base = (unsigned char *)dst[f].base;
    for (int k = 0; k < v.ncomp; k++) {
      HIPCHK(hipMemcpy(stage.data(), (const unsigned char *)v.c[k] + es * lo,
                       es * cnt, hipMemcpyDeviceToHost));
      for (size_t s2 = 0; s2 < cnt; s2++)
        std::memcpy(base + (int64_t)perm[s2] * dst[f].stride + es * k,
                    stage.data() + es * s2, es);
    }
  }
  return 0;
}

int swsph_num_cells(swsph_engine *e, int *ncells) {
  if (!e || !ncells) return fail(1, "null argument");
  *ncells = e->ncells;
  return 0;
}

int swsph_cell_count(swsph_engine *e, int cid, int *count) {
  if (!e || !count) return fail(1, "null argument");
  int rc = cell_ok(e, cid);
  if (rc) return rc;
  *count = e->cell_start[cid + 1] - e->cell_start[cid];
  return 0;
}

int swsph_sync(swsph_engine *e) {
  if (!e) return fail(1, "null argument");
  HIPCHK(hipStreamSynchronize(e->st));
  return 0;
}

int swsph_stream(swsph_engine *e, void **stream) {
  if (!e || !stream) return fail(1, "null argument");
  *stream = (void *)e->st;
  return 0;
}

// ---- per-task entry points ------------------------------------------------
int swsph_init_cell(swsph_engine *e, int ci) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  int rc = cell_ok(e, ci);
  if (rc) return rc;
  if ((rc = ensure_records(e))) return rc;
  HIPCHK(sphk::launch_init_density(e->d, e->cell_start[ci],
                                   e->cell_start[ci + 1],
                                   e->props.max_active_bin, e->props.h_max,
                                   e->st));
  e->launches++;
  return 0;
}

#define TASK_ENTRY(name, kind, subset, both)                        \
  int name(swsph_engine *e, int ci, int cj) {                       \
    if (!e) return fail(1, "null argument");                        \
    std::lock_guard<std::mutex> lk(e->mtx);                         \
    return task(e, kind, subset, ci, cj, both);                     \
  }
#define SELF_ENTRY(name, kind, subset)                              \
  int name(swsph_engine *e, int ci) {                               \
    if (!e) return fail(1, "null argument");                        \
    std::lock_guard<std::mutex> lk(e->mtx);                         \
    return task(e, kind, subset, ci, ci, false);                    \
  }

// A bundle of released tasks in one launch (the scheduler -> device task
// export): the directed interactions of all tasks are merged per i-cell into
// a list of neighbour images, and every chunk of such an i-cell becomes one
// work item.  Equivalent to running the tasks one by one (the same
// interactions; FP32 sums in another order).
int swsph_init_cells(swsph_engine *e, const int *cells, int n) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  return init_cells_locked(e, cells, n);
}

int swsph_do_ghost_cells(swsph_engine *e, const int *cells, int n,
                         int *rounds) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  return ghost_cells_locked(e, cells, n, rounds);
}

int swsph_do_extra_ghost_cells(swsph_engine *e, const int *cells, int n) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  return cell_op_locked(e, cells, n, true);
}

int swsph_do_end_hydro_force_cells(swsph_engine *e, const int *cells, int n) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  return cell_op_locked(e, cells, n, false);
}

int swsph_cell_active(swsph_engine *e, uint8_t *act) {
  if (!e || !act) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->need_particles) return fail(1, "no particles for the current grid");
  EngineTasks dev(e);
  std::vector<unsigned char> a;
  if (int rc = dev.cell_active(a)) return rc;
  std::memcpy(act, a.data(), a.size());
  return 0;
}

int swsph_sched_pass(swsph_engine *e, swsph_sched_stats *sched,
                     swsph_pass_stats *stats) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  return fitting(e, stats,
                 [&]() { return sched_pass_locked(e, sched, stats); });
}

int swsph_set_step_sched(swsph_engine *e, int on) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  e->step_sched = on ? 1 : 0;
  return 0;
}

int swsph_run_tasks(swsph_engine *e, int loop, const swsph_task *tasks,
                    int ntasks) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  return run_tasks_locked(e, loop, tasks, ntasks);
}

SELF_ENTRY(swsph_doself1_density, LoopKind::Density, false)
TASK_ENTRY(swsph_dopair1_density, LoopKind::Density, false, true)
SELF_ENTRY(swsph_doself1_gradient, LoopKind::Gradient, false)
TASK_ENTRY(swsph_dopair1_gradient, LoopKind::Gradient, false, true)
SELF_ENTRY(swsph_doself2_force, LoopKind::Force, false)
TASK_ENTRY(swsph_dopair2_force, LoopKind::Force, false, true)

// The density of a subset of cell ci's particles (caller indices
// first + ind[k]) against cell cj (ci itself for the self form).
int subset_task(swsph_engine *e, int ci, int cj, int64_t first,
                const int *ind, int count) {
  int rc;
  if ((rc = cell_ok(e, ci)) || (rc = cell_ok(e, cj))) return rc;
  if (count < 0 || (count > 0 && !ind) || first < 0)
    return fail(1, "bad subset (first %lld, count %d)", (long long)first,
                count);
  for (int k = 0; k < count; k++)
    if (ind[k] < 0 || first + ind[k] >= e->n)
      return fail(1, "subset entry %d: particle %lld out of range [0, %lld)",
                  k, (long long)(first + ind[k]), (long long)e->n);
  if ((rc = ensure_subset(e, count))) return rc;
  const int lo = e->cell_start[ci], hi = e->cell_start[ci + 1];
  if (count > 0)
    HIPCHK(hipMemcpyAsync(e->d_subind, ind, sizeof(int) * (size_t)count,
                          hipMemcpyHostToDevice, e->st));
  HIPCHK(sphk::launch_mark_subset(e->d_subflag, e->d_inv, (long long)first,
                                  e->d_subind, count, lo, hi, e->d_suberr,
                                  e->st));
  e->launches += 3;
  unsigned int bad = 0;
  HIPCHK(hipMemcpyAsync(&bad, e->d_suberr, sizeof(bad), hipMemcpyDeviceToHost,
                        e->st));
  // (also the end of the host read of ind)
  HIPCHK(hipStreamSynchronize(e->st));
  if (bad)
    return fail(1, "%u particles of the subset are not in cell %d", bad, ci);
  return task(e, LoopKind::Density, true, ci, cj, false, e->d_subflag);
}

int swsph_doself_subset_density(swsph_engine *e, int ci, int64_t first,
                                const int *ind, int count) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  return subset_task(e, ci, ci, first, ind, count);
}

int swsph_dopair_subset_density(swsph_engine *e, int ci, int64_t first,
                                const int *ind, int count, int cj) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  return subset_task(e, ci, cj, first, ind, count);
}

int swsph_do_ghost(swsph_engine *e, int ci, int offset, int ntasks) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  int rc = cell_ok(e, ci);
  if (rc) return rc;
  if (ntasks < 1 || offset < 0 || offset >= ntasks)
    return fail(1, "ghost share %d of %d", offset, ntasks);
  int rounds = 0;
  if ((rc = ghost_cell(e, ci, &rounds, offset, ntasks))) return rc;
  // a support past the cell width would truncate the neighbour sets of the
  // tasks that follow: status 4 (the caller rebuilds its cells)
  return check_fits(e);
}

int swsph_do_extra_ghost(swsph_engine *e, int ci) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  int rc = cell_ok(e, ci);
  if (rc) return rc;
  return extra_ghost_cell(e, ci);
}

int swsph_do_end_hydro_force(swsph_engine *e, int ci) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  int rc = cell_ok(e, ci);
  if (rc) return rc;
  HIPCHK(sphk::launch_end_force(e->d, e->cell_start[ci], e->cell_start[ci + 1],
                                e->props.max_active_bin, e->st));
  e->launches++;
  return 0;
}

// ---- whole passes -----------------------------------------------------------
int swsph_pass_plain(swsph_engine *e, swsph_pass_stats *stats) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->need_particles)
    return fail(1, "the grid changed: call swsph_set_particles first");
  if (e->n == 0) return empty_pass(stats);
  int rc = 0;
  const int64_t l0 = e->launches;
  std::vector<int> up;
  e->density_work = 0;
  if ((rc = ensure_records(e))) return rc;
  e->last_pass_mab = e->props.max_active_bin;
  HIPCHK(sphk::launch_zero_u64(e->d_redo + 3, kGhostWorkSlots, e->st));
  HIPCHK(sphk::launch_zero_u32(e->d_hg2_all, 1, e->st));
  for (int c = 0; c < e->ncells; c++) {
    HIPCHK(sphk::launch_init_density(e->d, e->cell_start[c], e->cell_start[c + 1],
                                     e->props.max_active_bin, e->props.h_max,
                                     e->st));
    e->launches++;
  }
  auto loop_tasks = [&](LoopKind k) -> int {
    int r;
    for (int c = 0; c < e->ncells; c++) {
      if ((r = task(e, k, false, c, c, false))) return r;
      upper_neighbours(e, c, up);
      for (int cj : up)
        if ((r = task(e, k, false, c, cj, true))) return r;
    }
    return 0;
  };
  if ((rc = loop_tasks(LoopKind::Density))) return rc;
  int rounds = 0;
  for (int c = 0; c < e->ncells; c++)
    if ((rc = ghost_cell(e, c, &rounds))) return rc;
  if ((rc = loop_tasks(LoopKind::Gradient))) return rc;
  for (int c = 0; c < e->ncells; c++)
    if ((rc = extra_ghost_cell(e, c))) return rc;
  if ((rc = loop_tasks(LoopKind::Force))) return rc;
  for (int c = 0; c < e->ncells; c++) {
    HIPCHK(sphk::launch_end_force(e->d, e->cell_start[c], e->cell_start[c + 1],
                                  e->props.max_active_bin, e->st));
    e->launches++;
  }
  if (stats) {
    stats->ghost_rounds = rounds;
    stats->launches = e->launches - l0;
    stats->list_rebuilt = 0;
    stats->list_fallback = 0;
    stats->n_density_all = e->density_work;
  }
  if ((rc = count_into(e, stats))) return rc;
  return check_support(e);
}


int swsph_pass_batched(swsph_engine *e, swsph_pass_stats *stats) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  return fitting(e, stats, [&]() { return pass_batched_locked(e, stats); });
}

int swsph_set_regrid(swsph_engine *e, int on) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  e->auto_regrid = on != 0;
  return 0;
}

int swsph_get_grid(swsph_engine *e, int ncell[3]) {
  if (!e || !ncell) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  for (int k = 0; k < 3; k++) ncell[k] = e->grid.nc[k];
  return 0;
}

int swsph_pass_batched_noalloc(swsph_engine *e, swsph_pass_stats *stats) {
  if (!e) return fail(1, "null engine");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->n > kNoAllocMaxParts || e->ncells > kNoAllocMaxCells)
    return fail(1, "the allocation-free pass takes at most %lld particles on "
                "%d cells (%lld particles on %d cells given)",
                (long long)kNoAllocMaxParts, kNoAllocMaxCells,
                (long long)e->n, e->ncells);
  if (e->use_lists)
    return fail(1, "neighbour lists allocate their storage: disable them "
                "(swsph_set_lists) for the allocation-free pass");
  if (e->timing)
    return fail(1, "kernel timing creates events: disable it "
                "(swsph_set_timing) for the allocation-free pass");
  TileScratch ts{};
  size_t tile_cap = 0;
  unsigned long long *redo = nullptr, *cnt = nullptr;
  HIPCHK(sphk::noalloc_scratch(e->device, e->n, e->ncells, &ts, &tile_cap,
                               &redo, &cnt));
  // The pass runs on the scratch in place of the engine's own work queue
  // and counters; the all-active tile table it leaves there is not the
  // engine's (tiles_c = 0 before and after).
  std::swap(e->ts, ts);
  std::swap(e->tile_cap, tile_cap);
  std::swap(e->d_redo, redo);
  std::swap(e->d_cnt, cnt);
  e->tiles_c = 0;
  e->no_alloc = true;
  const int rc = pass_batched_locked(e, stats);
  e->no_alloc = false;
  e->tiles_c = 0;
  std::swap(e->ts, ts);
  std::swap(e->tile_cap, tile_cap);
  std::swap(e->d_redo, redo);
  std::swap(e->d_cnt, cnt);
  return rc;
}

static int pass_graph_locked(swsph_engine *e, swsph_pass_stats *stats);

int swsph_pass_graph(swsph_engine *e, swsph_pass_stats *stats) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  return fitting(e, stats, [&]() { return pass_graph_locked(e, stats); });
}

static int pass_graph_locked(swsph_engine *e, swsph_pass_stats *stats) {
  int rc = 0;
  if (e->use_lists)
    return fail(1, "graph replay needs the neighbour lists disabled");
  if (e->need_particles)
    return fail(1, "the grid changed: call swsph_set_particles first");
  if (e->n == 0) return empty_pass(stats);
  const int64_t l0 = e->launches;
  e->list_rebuilt = 0;
  e->list_fallback = 0;
  e->density_work = 0;
  if ((rc = ensure_records(e))) return rc;
  e->last_pass_mab = e->props.max_active_bin;
  // Re-runs captured: as many as the last pass needed.
  // (at most kHistRounds - 1: one redo count per captured ghost round)
  const int reruns = std::min(
      swsph_engine::kHistRounds - 1,
      std::max(1, std::min(e->props.max_iter, e->last_rounds) - 1));
  if (!e->g_density || e->graph_gen != e->gen || e->graph_reruns != reruns) {
    drop_graphs(e);
    // chunk tables must exist before capture (they allocate)
    for (int k = 0; k < 3; k++)
      if ((rc = ensure_tables(e, k))) return rc;
    // the all-active tile count of each tile capacity in use
    e->cap_tiles.clear();
    if (e->schedule == 1) {
      const LoopArgs la = loop_args(e);
      for (int k = 0; k < 3; k++) {
        const int C = e->loop_c[k];
        if (e->cap_tiles.count(C)) continue;
        HIPCHK(sphk::launch_build_tiles(e->d, la, (int)e->n, e->ncells, C, 0,
                                        nullptr, e->ts, e->st));
        HIPCHK(hipMemcpyAsync(e->h_ntiles, e->ts.ntiles, sizeof(int),
                              hipMemcpyDeviceToHost, e->st));
        HIPCHK(hipStreamSynchronize(e->st));
        e->cap_tiles[C] = *e->h_ntiles;
      }
      e->tiles_c = 0;
    }
    // after each ghost round its redo count is copied aside, so that a
    // replay tells how many of the captured rounds had work
    auto ghost_round = [&](int k) -> int {
      if (int r = phase_ghost(e)) return r;
      HIPCHK(hipMemcpyAsync(e->d_hist + k, e->d_redo,
                            sizeof(unsigned long long),
                            hipMemcpyDeviceToDevice, e->st));
      return 0;
    };
    if ((rc = capture(e, &e->g_density, [&]() -> int {
           int r = phase_start(e);
           for (int k = 0; k < reruns && !r; k++) {
             if ((r = ghost_round(k))) break;
             r = phase_rerun(e, 0, 0);
           }
           return r ? r : ghost_round(reruns);
         })))
      return rc;
    if ((rc = capture(e, &e->g_finish, [&]() -> int { return phase_finish(e); })))
      return rc;
    e->graph_gen = e->gen;
    e->graph_reruns = reruns;
  }
  HIPCHK(hipGraphLaunch(e->g_density, e->st));
  HIPCHK(hipMemcpyAsync(e->h_hist, e->d_hist,
                        (reruns + 1) * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, e->st));
  int redo = 0;
  if ((rc = read_redo(e, &redo))) return rc;  // waits for the copy too
  // rounds this replay needed: up to the first ghost that flagged none
  int needed = reruns + 1;
  for (int k = 0; k <= reruns; k++)
    if (e->h_hist[k] == 0) {
      needed = k + 1;
      break;
    }
  int rounds = reruns + 1;
  // more rounds than captured: continue eagerly (and capture more next time)
  while (redo > 0 && rounds < e->props.max_iter) {
    if ((rc = phase_rerun(e, redo, 0)) || (rc = phase_ghost(e)) ||
        (rc = read_redo(e, &redo)))
      return rc;
    rounds++;
  }
  if (redo > 0)
    return fail(3, "Smoothing length failed to converge on %d particles", redo);
  // More rounds than needed were captured: after 3 such replays in a row
  // the next pass captures fewer (a count that changes from pass to pass
  // does not recapture every time).
  if (rounds == reruns + 1 && needed < rounds) {
    if (++e->shrink_votes >= 3) {
      rounds = needed;
      e->shrink_votes = 0;
    }
  } else {
    e->shrink_votes = 0;
  }
  e->last_rounds = rounds;
  HIPCHK(hipGraphLaunch(e->g_finish, e->st));
  return finish_stats(e, stats, rounds, l0);
}

int swsph_save_state(swsph_engine *e) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  std::vector<float *> src = io_fields(e->d);
  const size_t n = (size_t)(e->n ? e->n : 1);
  {
    int rc = ensure_copies(e, e->saved, e->saved_cap, src.size(), n);
    if (rc) return rc;
  }
  sphk::CopySet cs{};
  cs.m = (int)src.size();
  if (cs.m > sphk::kMaxCopies) return fail(2, "too many state fields");
  for (int k = 0; k < cs.m; k++) {
    cs.src[k] = src[k];
    cs.dst[k] = e->saved[k];
  }
  HIPCHK(sphk::launch_copy_set(cs, (int)n, e->st));
  e->saved_valid = true;
  e->saved_epoch = e->fields_epoch;
  return 0;
}

int swsph_restore_state(swsph_engine *e) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->saved.empty() || !e->saved_valid)
    return fail(1, "no saved state for the current particles");
  std::vector<float *> dst = io_fields(e->d);
  const size_t n = (size_t)(e->n ? e->n : 1);
  if (e->saved_cap < n)
    return fail(1, "no saved state for the current particles");
  sphk::CopySet cs{};
  cs.m = (int)dst.size();
  if (cs.m > sphk::kMaxCopies) return fail(2, "too many state fields");
  for (int k = 0; k < cs.m; k++) {
    cs.src[k] = e->saved[k];
    cs.dst[k] = dst[k];
  }
  HIPCHK(sphk::launch_copy_set(cs, (int)n, e->st));
  if (e->saved_epoch != e->fields_epoch) {
    invalidate_records(e);  // fields changed outside passes since the save
    e->fields_epoch++;
  } else {
    e->stale_mab = e->last_pass_mab;
  }
  return 0;
}

int swsph_set_chunk(swsph_engine *e, int which, int C) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  e->gen++;  // captured graphs no longer match
  if (which < 0 || which > 2) return fail(1, "loop index %d out of range", which);
  if (C != 8 && C != 16 && C != 32 && C != 64)
    return fail(1, "chunk capacity %d not in {8, 16, 32, 64}", C);
  e->loop_c[which] = C;
  e->loop_cj[which] = C;
  prune_tables(e);
  if (!e->cell_start.empty()) {
    int rc = 0;
    if (!chunk_table(e, C, &rc)) return rc;
  }
  return 0;
}

int swsph_set_tile_shape(swsph_engine *e, int which, int CI, int CJ) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  e->gen++;  // captured graphs no longer match
  if (which < 0 || which > 2) return fail(1, "loop index %d out of range", which);
  const bool ok = (CJ == CI && (CI == 8 || CI == 16 || CI == 32 || CI == 64)) ||
                  (CJ * 2 == CI && (CI == 16 || CI == 32)) ||
                  (CI == 4 && (CJ == 8 || CJ == 16 || CJ == 32)) ||
                  (CI == 8 && CJ == 16) || (CI == 16 && CJ == 32);
  if (!ok) return fail(1, "tile shape (%d, %d) not supported", CI, CJ);
  e->loop_c[which] = CI;
  e->loop_cj[which] = CJ;
  prune_tables(e);
  if (!e->cell_start.empty()) return ensure_tables(e, which);
  return 0;
}

int swsph_set_bin_order(swsph_engine *e, int on) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->bin_order == (on != 0)) return 0;
  e->bin_order = on != 0;
  e->gen++;  // captured graphs no longer match
  // resident particles are re-binned in the new order now
  if (e->n > 0 && !e->need_particles) {
    e->tiles_c = 0;
    return rebin(e);
  }
  return 0;
}

int swsph_set_sparse_shape(swsph_engine *e, int CI, int CJ, int per_cell) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (per_cell < 0) return fail(1, "per_cell %d is negative", per_cell);
  if (per_cell > 0) {
    const bool ok = (CJ == CI && (CI == 8 || CI == 16 || CI == 32 || CI == 64)) ||
                    (CJ * 2 == CI && (CI == 16 || CI == 32)) ||
                    (CI == 4 && (CJ == 8 || CJ == 16 || CJ == 32)) ||
                    (CI == 8 && CJ == 16);
    if (!ok) return fail(1, "tile shape (%d, %d) not supported", CI, CJ);
    e->sparse_ci = CI;
    e->sparse_cj = CJ;
  }
  e->sparse_per_cell = per_cell;
  return 0;
}

int swsph_get_tile_shape(swsph_engine *e, int which, int *CI, int *CJ) {
  if (!e || !CI || !CJ) return fail(1, "null argument");
  if (which < 0 || which > 2) return fail(1, "loop index %d out of range", which);
  *CI = e->loop_c[which];
  *CJ = e->loop_cj[which];
  return 0;
}

int swsph_set_schedule(swsph_engine *e, int mode, int waves) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  e->gen++;  // captured graphs no longer match
  if (mode < 0 || mode > 2) return fail(1, "schedule %d not in {0, 1, 2}", mode);
  if (waves < 0) return fail(1, "negative wave count");
  e->schedule = mode;
  e->queue_waves = waves;
  prune_tables(e);
  return 0;
}

int swsph_set_lists(swsph_engine *e, int enable, float skin, int slots) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  e->gen++;  // captured graphs no longer match
  if (!(skin >= 1.f && skin <= 2.f))
    return fail(1, "list skin %g outside [1, 2]", (double)skin);
  if (slots < 1 || slots > 4096)
    return fail(1, "list slots %d outside [1, 4096]", slots);
  e->use_lists = enable != 0;
  e->list_skin = skin;
  if (slots != e->list_slots) {
    e->list_slots = slots;
    e->nl_elems = 0;  // reallocated on next use
  }
  e->nl.lam2 = skin * skin;
  e->nl.Ks = slots;
  return 0;
}

int swsph_get_chunk(swsph_engine *e, int which, int *C) {
  if (!e || !C) return fail(1, "null argument");
  if (which < 0 || which > 2) return fail(1, "loop index %d out of range", which);
  *C = e->loop_c[which];
  return 0;
}

int swsph_count_interactions(swsph_engine *e, int64_t out3[3]) {
  if (!e || !out3) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  swsph_pass_stats s{};
  int rc = count_into(e, &s);
  if (rc) return rc;
  out3[0] = s.n_density;
  out3[1] = s.n_gradient;
  out3[2] = s.n_force;
  return 0;
}

int swsph_set_timing(swsph_engine *e, int on) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  e->timing = on != 0;
  for (int k = 0; k < 3; k++) {
    e->t_ms[k] = 0.0;
    e->t_launch[k] = 0;
  }
  e->ev_pending.clear();
  e->ev_used = 0;
  return 0;
}

int swsph_get_timing(swsph_engine *e, double ms[3], int64_t launches[3]) {
  if (!e || !ms || !launches) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  int rc = collect_timing(e);
  if (rc) return rc;
  for (int k = 0; k < 3; k++) {
    ms[k] = e->t_ms[k];
    launches[k] = e->t_launch[k];
  }
  return 0;
}

int swsph_autotune(swsph_engine *e, int repeats, swsph_tune_result *out) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->n == 0) return fail(1, "no particles set");
  if (repeats < 1 || repeats > 100) return fail(1, "repeats %d outside [1, 100]", repeats);
  int rc = 0;
  if ((rc = copy_state(e, e->tune_saved, true))) return rc;
  const int sched0 = e->schedule, waves0 = e->queue_waves;
  const int c0[3] = {e->loop_c[0], e->loop_c[1], e->loop_c[2]};
  const int cj0[3] = {e->loop_cj[0], e->loop_cj[1], e->loop_cj[2]};
  const bool timing0 = e->timing;
  int trials = 0;
  // Times `repeats` passes of the current settings: per-loop kernel time
  // and wall time per pass (ms).
  auto time_config = [&](double ms[3], double *wall) -> int {
    int r = 0;
    e->gen++;
    // warm-up pass (builds the chunk tables of these capacities)
    if ((r = copy_state(e, e->tune_saved, false)) ||
        (r = pass_batched_locked(e, nullptr)))
      return r;
    e->timing = true;
    for (int k = 0; k < 3; k++) e->t_ms[k] = 0.0, e->t_launch[k] = 0;
    e->ev_pending.clear();
    e->ev_used = 0;
    HIPCHK(hipStreamSynchronize(e->st));
    const auto t0 = std::chrono::steady_clock::now();
    for (int q = 0; q < repeats && !r; q++)
      if (!(r = copy_state(e, e->tune_saved, false)))
        r = pass_batched_locked(e, nullptr);
    if (r) return r;
    if ((r = collect_timing(e))) return r;
    const auto t1 = std::chrono::steady_clock::now();
    *wall = std::chrono::duration<double, std::milli>(t1 - t0).count() / repeats;
    for (int k = 0; k < 3; k++) ms[k] = e->t_ms[k] / repeats;
    e->timing = false;
    trials++;
    return 0;
  };
  // With lists the density loop builds them in tiles of kListChunk.
  const int caps[3] = {8, 16, 32};
  double best_total = 1e300;
  int best_s = sched0, best_c[3] = {c0[0], c0[1], c0[2]};
  int best_cj[3] = {cj0[0], cj0[1], cj0[2]};
  double best_ms[3] = {0, 0, 0}, best_pass = 0;
  double tab[3][3][3];  // [schedule][capacity][loop] ms per pass
  double wall[3][3];
  for (int sc = 0; sc < 3 && !rc; sc++) {
    for (int ci = 0; ci < 3 && !rc; ci++) {
      e->schedule = sc;
      e->queue_waves = 0;
      for (int k = 0; k < 3; k++) e->loop_c[k] = e->loop_cj[k] = caps[ci];
      if (e->use_lists) e->loop_c[0] = e->loop_cj[0] = kListChunk;
      rc = time_config(tab[sc][ci], &wall[sc][ci]);
    }
  }
  if (!rc) {
    // best capacity per loop under each schedule, then the best schedule
    for (int sc = 0; sc < 3; sc++) {
      double tot = 0.0;
      int cs[3];
      double ms[3];
      for (int k = 0; k < 3; k++) {
        int bi = 0;
        for (int ci = 1; ci < 3; ci++)
          if (tab[sc][ci][k] < tab[sc][bi][k]) bi = ci;
        cs[k] = caps[bi];
        ms[k] = tab[sc][bi][k];
        tot += ms[k];
      }
      if (e->use_lists) cs[0] = kListChunk;
      if (tot < best_total) {
        best_total = tot;
        best_s = sc;
        for (int k = 0; k < 3; k++)
          best_c[k] = best_cj[k] = cs[k], best_ms[k] = ms[k];
        best_pass = wall[sc][1];
      }
    }
  }
  if (!rc && best_s >= 1) {
    // Under the work queue: candidate chunks of half the tile size, then
    // the narrow tiles (4, 32) and (8, 16) and the sparse tiles (4, 8) (a
    // few active particles per cell), each loop keeping its fastest shape.
    // The trials time eager passes (one wave per tile); a captured graph
    // sizes its grids by the tile count of its configuration, so the narrow
    // tiles cost no idle waves there either.
    e->schedule = best_s;
    e->queue_waves = 0;
    int best_ci[3] = {best_c[0], best_c[1], best_c[2]};
    for (int trial = 0; trial < 4 && !rc; trial++) {
      bool any = false;
      for (int k = 0; k < 3; k++) {
        int ci = best_c[k], cj = best_c[k];
        if (!(e->use_lists && k == 0)) {
          if (trial == 0 && (best_c[k] == 16 || best_c[k] == 32)) cj = ci / 2;
          if (trial == 1) ci = 4, cj = 32;
          if (trial == 2) ci = 8, cj = 16;
          if (trial == 3) ci = kMinTile, cj = kSparseSub;
        }
        e->loop_c[k] = ci;
        e->loop_cj[k] = cj;
        any = any || ci != best_c[k] || cj != best_c[k];
      }
      double ms[3], w = 0;
      if (!any || (rc = time_config(ms, &w))) continue;
      for (int k = 0; k < 3; k++)
        if ((e->loop_c[k] != best_c[k] || e->loop_cj[k] != best_c[k]) &&
            ms[k] < best_ms[k]) {
          best_ci[k] = e->loop_c[k];
          best_cj[k] = e->loop_cj[k];
          best_ms[k] = ms[k];
        }
    }
    for (int k = 0; k < 3; k++) best_c[k] = best_ci[k];
  }
  // apply the choice (or restore the settings after a failure) and the state
  e->schedule = rc ? sched0 : best_s;
  e->queue_waves = rc ? waves0 : 0;
  for (int k = 0; k < 3; k++) {
    e->loop_c[k] = rc ? c0[k] : best_c[k];
    e->loop_cj[k] = rc ? cj0[k] : best_cj[k];
  }
  prune_tables(e);
  e->gen++;
  e->timing = timing0;
  for (int k = 0; k < 3; k++) e->t_ms[k] = 0.0, e->t_launch[k] = 0;
  e->ev_pending.clear();
  e->ev_used = 0;
  const int rc2 = copy_state(e, e->tune_saved, false);
  if (rc) return rc;
  if (rc2) return rc2;
  HIPCHK(hipStreamSynchronize(e->st));
  if (out) {
    out->schedule = best_s;
    for (int k = 0; k < 3; k++) {
      out->chunk[k] = best_c[k];
      out->chunk_j[k] = best_cj[k];
      out->loop_ms[k] = best_ms[k];
    }
    out->pass_ms = best_pass;
    out->trials = trials;
  }
  return 0;
}

// ---- time integration -------------------------------------------------------

int swsph_set_integration(swsph_engine *e, const swsph_integration *p) {
  if (!e || !p) return fail(1, "null argument");
  if (!(p->cfl > 0.f) || !(p->dt_max > 0.0) || p->dt_min < 0.0 ||
      !(p->log_max_h_change > 0.f) || p->min_u < 0.f)
    return fail(1, "invalid integration parameters");
  std::lock_guard<std::mutex> lk(e->mtx);
  e->integ = *p;
  return 0;
}

int swsph_do_drift_part(swsph_engine *e, int ci, int64_t ti_old,
                        int64_t ti_current) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->need_particles) return fail(1, "no particles for the current grid");
  e->gen++;  // positions moved: captured graphs hold stale bounds
  return drift_cells(e, ci, ti_old, ti_current);
}

int swsph_do_kick1(swsph_engine *e, int ci, int64_t ti_current) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->need_particles) return fail(1, "no particles for the current grid");
  return kick_cells(e, ci, ti_current, 1);
}

int swsph_do_kick2(swsph_engine *e, int ci, int64_t ti_current) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->need_particles) return fail(1, "no particles for the current grid");
  return kick_cells(e, ci, ti_current, 2);
}

int swsph_do_timestep(swsph_engine *e, int ci, int64_t ti_current,
                      int64_t *ti_end_min) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->need_particles) return fail(1, "no particles for the current grid");
  if (!(e->props.time_base > 0.0)) return fail(1, "time_base must be positive");
  return timestep_cells(e, ci, ti_current, ti_end_min, nullptr);
}

int swsph_rebuild(swsph_engine *e) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->need_particles) return fail(1, "no particles for the current grid");
  int rc = rebin(e);
  if (rc) return rc;
  return check_fits(e);
}

int swsph_engine_step(swsph_engine *e, swsph_step_info *info) {
  if (!e) return fail(1, "null argument");
  std::lock_guard<std::mutex> lk(e->mtx);
  if (e->need_particles)
    return fail(1, "the grid changed: call swsph_set_particles first");
  if (!(e->props.time_base > 0.0))
    return fail(1, "time_base must be positive for time integration");
  if (e->n == 0) return fail(1, "no particles to integrate");
  int rc = 0, rebuilt = 0;
  const int64_t l0 = e->launches;
  const bool first = !e->started;
  if (first) {
    e->ti_current = 0;
  } else {
    const int64_t ti_old = e->ti_current;
    e->ti_current = e->ti_end_min;
    e->gen++;
    if ((rc = drift_cells(e, -1, ti_old, e->ti_current))) return rc;
    double H, dm;
    if ((rc = read_fit(e, &H, &dm))) return rc;
    if (!fits(e, H, dm)) {
      if ((rc = rebin(e))) return rc;
      rebuilt = 1;
    }
  }
  const float dx_max = e->dmax;
  set_active_bin(e, max_active_bin_at(e->ti_current));
  // narrow tiles when the active particles are few per cell, and list
  // sweeps when they are under half of all
  bool sparse = false;
  if (e->schedule >= 1) {
    HIPCHK(sphk::launch_count_active(e->d, (int)e->n, e->props.max_active_bin,
                                     e->d_err + 2, e->st));
    HIPCHK(hipMemcpyAsync(e->h_err + 2, e->d_err + 2, sizeof(unsigned int),
                          hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    const double nact = (double)e->h_err[2];
    sparse = e->sparse_per_cell > 0 &&
             nact < (double)e->sparse_per_cell * (double)e->ncells;
    e->sweep_list = 2.0 * nact < (double)e->n;
  }
  const int c0[3] = {e->loop_c[0], e->loop_c[1], e->loop_c[2]};
  const int cj0[3] = {e->loop_cj[0], e->loop_cj[1], e->loop_cj[2]};
  if (sparse) {
    for (int k = 0; k < 3; k++) {
      e->loop_c[k] = e->sparse_ci;
      e->loop_cj[k] = e->sparse_cj;
    }
    e->tiles_c = 0;
  }
  swsph_pass_stats ps{};
  auto pass = [&]() {
    return e->step_sched ? sched_pass_locked(e, nullptr, &ps)
                         : pass_batched_locked(e, &ps);
  };
  // the smoothing lengths may outgrow the slack the drift left (re-bin) or
  // the cells themselves (regrid): the pass is then repeated
  const uint64_t pg0 = e->perm_gen;
  rc = fitting(e, &ps, pass);
  if (e->perm_gen != pg0) rebuilt = 1;
  e->sweep_list = false;
  if (sparse) {
    for (int k = 0; k < 3; k++) {
      e->loop_c[k] = c0[k];
      e->loop_cj[k] = cj0[k];
    }
    e->tiles_c = 0;
  }
  if (rc) return rc;
  if (!first && (rc = kick_cells(e, -1, e->ti_current, 2))) return rc;
  int64_t tiend = 0, nact = 0;
  if ((rc = timestep_cells(e, -1, e->ti_current, &tiend, &nact))) return rc;
  if ((rc = kick_cells(e, -1, e->ti_current, 1))) return rc;
  HIPCHK(hipStreamSynchronize(e->st));
  if (tiend <= e->ti_current)
    return fail(3, "the next synchronisation point does not advance");
  e->ti_end_min = tiend;
  e->started = true;
  if (info) {
    info->sparse = sparse ? 1 : 0;
    info->ti_current = e->ti_current;
    info->ti_end_min = tiend;
    info->max_active_bin = e->props.max_active_bin;
    info->rebuilt = rebuilt;
    info->n_active = nact;
    info->dx_max = dx_max;
    info->launches = e->launches - l0;
    info->pass = ps;
  }
  return 0;
}

}  // extern "C"
