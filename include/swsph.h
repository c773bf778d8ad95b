/*
 * swsph.h -- C ABI of the MI355X SPHENIX neighbour pass (libswsph.so).
 *
 * Plain pointers and sizes only.  Every call returns 0 on success and a
 * non-zero status on error; swsph_last_error() returns the message of the
 * last failure on the calling thread.  A SWIFT-side shim turns a non-zero
 * status into error() (src/error.h:60-80), the reference's fail-stop
 * convention.
 *
 * Model: an engine owns a device-resident copy of the gas particles
 * (structure of arrays in HBM, sorted by top-level cell) and the top-level
 * cell grid.  Cells are addressed by their index
 * cid = (cx * ncell[1] + cy) * ncell[2] + cz on that grid.  The per-task
 * entry points below are the drop-in for SWIFT's task functions; each one
 * enqueues its work on the engine's stream and returns (the ghost waits for
 * its own convergence test, as the reference's does).  Calls on one engine
 * are serialised internally.
 *
 * Scope: SPHENIX, Wendland C2, 3D, gamma = 5/3, no cosmology (a = 1, H = 0),
 * periodic box with >= 3 top-level cells per axis and kernel support
 * H = gamma_K h <= top-level cell width (the whole-pass forms regrid when
 * it grows past it, see swsph_set_regrid).
 */
#ifndef SWSPH_API_H_
#define SWSPH_API_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct swsph_engine swsph_engine;

/* Run-time parameters read on this path: struct hydro_props
 * (src/hydro_properties.h; defaults src/hydro_properties.c:348-390 and
 * src/hydro/SPHENIX/hydro_parameters.h:59-101), the engine's time_base and
 * max_active_bin (src/engine.h), and the space geometry (src/space.h). */
typedef struct swsph_props {
  double dim[3];          /* periodic box size */
  int ncell[3];           /* top-level cells per axis */
  float eta;              /* SPH:resolution_eta */
  float h_tolerance;      /* SPH:h_tolerance */
  float h_max, h_min;     /* SPH:h_max, h_min */
  int max_iter;           /* SPH:max_ghost_iterations */
  float alpha_max, alpha_min, visc_length;          /* viscosity */
  float diff_beta, diff_alpha_max, diff_alpha_min;  /* diffusion */
  double time_base;       /* dt of bin b = 2^(b+1) * time_base */
  int max_active_bin;     /* part_is_active: time_bin <= max_active_bin */
} swsph_props;

/* Host particle arrays in the caller's order.  Vectors are [3n] interleaved
 * like struct part.x[3] / v[3] (src/hydro/SPHENIX/hydro_part.h:105-324).
 * The (io) fields are the state an inactive particle brings into the pass
 * and are written back for active particles by swsph_get_field. */
typedef struct swsph_host_parts {
  int64_t n;
  const double *x;
  const float *v;
  const float *m;
  const float *u;
  const int8_t *time_bin;
  const float *h;           /* (io) smoothing length */
  const float *rho;         /* (io) */
  const float *pressure;    /* (io) force.pressure */
  const float *soundspeed;  /* (io) force.soundspeed */
  const float *f;           /* (io) force.f */
  const float *balsara;     /* (io) force.balsara */
  const float *alpha_visc;  /* (io) viscosity.alpha */
  const float *alpha_diff;  /* (io) diffusion.alpha */
  const float *div_v_prev;  /* (io) viscosity.div_v_previous_step */
} swsph_host_parts;

/* Field identifiers for swsph_get_field (host buffer in the caller's order;
 * float unless noted, vectors [3n]). */
enum swsph_field {
  SWSPH_H = 0, SWSPH_RHO, SWSPH_RHO_DH, SWSPH_WCOUNT, SWSPH_WCOUNT_DH,
  SWSPH_DIV_V, SWSPH_ROT_V, SWSPH_PRESSURE, SWSPH_SOUNDSPEED, SWSPH_F,
  SWSPH_BALSARA, SWSPH_V_SIG, SWSPH_LAPLACE_U, SWSPH_ALPHA_VISC_MAX_NGB,
  SWSPH_ALPHA_VISC, SWSPH_ALPHA_DIFF, SWSPH_DIV_V_DT, SWSPH_DIV_V_PREV,
  SWSPH_A_HYDRO, SWSPH_U_DT, SWSPH_H_DT,
  SWSPH_N_DENSITY,          /* int32: neighbours j != i, r < H_i */
  SWSPH_N_GRADIENT,         /* int32 */
  SWSPH_N_FORCE,            /* int32: neighbours j != i, r < max(H_i, H_j) */
  SWSPH_MIN_NGB_TIME_BIN,   /* int8 */
  SWSPH_NUM_FIELDS,         /* count of the pass outputs above */
  /* particle state advanced by the time integration below */
  SWSPH_X = 32,             /* double [3n]: positions (part.x) */
  SWSPH_V,                  /* [3n] predicted velocities (part.v) */
  SWSPH_V_FULL,             /* [3n] xpart.v_full */
  SWSPH_U,                  /* predicted internal energy (part.u) */
  SWSPH_U_FULL,             /* xpart.u_full */
  SWSPH_TIME_BIN            /* int8 */
};

typedef struct swsph_pass_stats {
  int ghost_rounds;              /* density evaluations of the slowest particle */
  int64_t n_density;             /* sum of N_density over active particles */
  int64_t n_gradient;
  int64_t n_force;
  int64_t launches;              /* kernel launches issued by the pass */
  int64_t list_rebuilt;          /* batched: particles re-listed by re-runs */
  int64_t list_fallback;         /* batched: gradient via the chunk loop */
  int64_t n_density_all;         /* batched: density interactions evaluated
                                    over all ghost rounds (work done) */
  int64_t regrids;               /* whole passes: top-level regrids the pass
                                    needed (kernel support outgrew a cell) */
} swsph_pass_stats;

const char *swsph_last_error(void);

/* Engine life cycle. */
int swsph_engine_create(const swsph_props *props, swsph_engine **out);
int swsph_engine_destroy(swsph_engine *e);
int swsph_set_props(swsph_engine *e, const swsph_props *props);

/* Uploads the particles (arrays in the caller's order, positions inside the
 * box) and bins them into the top-level grid on the device: sorted by cell
 * and by a Morton code inside each cell.  Replaces space_rebuild's particle
 * sort for this path.  Starts a new time line for swsph_engine_step. */
int swsph_set_particles(swsph_engine *e, const swsph_host_parts *p);
int swsph_get_field(swsph_engine *e, int field, void *host_out);
/* nf fields at once (as swsph_get_field each), with one wait at the end:
 * with page-locked outputs (hipHostMalloc) the copies are asynchronous
 * DMA behind each other on the engine's stream. */
int swsph_get_fields(swsph_engine *e, int nf, const int *fields,
                     void *const *outs);
int swsph_num_cells(swsph_engine *e, int *ncells);
int swsph_cell_count(swsph_engine *e, int cid, int *count);
int swsph_sync(swsph_engine *e);

/* ---- drop-in task entry points ------------------------------------------
 * Each replaces the task function runner_main.c calls for a self or pair
 * task of the hydro loops (src/runner_main.c:215-300): runner_dosub_self1_-
 * density(r, ci, below_h_max = 0, 1) and its siblings
 * (src/runner_doiact_functions_hydro.h:2933-3092), on a top-level cell
 * (pair).  The reference's call recurses through the cell's tree and
 * processes every particle at the tree level its h selects: the depth
 * windows limit_h_min / limit_h_max of DOSELF1_BRANCH / DOPAIR1_BRANCH
 * (src/runner_doiact_hydro.h:176-184) split the particles between levels
 * so that each interaction is done once.  The device holds no sub-cells:
 * one call here does the interactions of every level of the top-level cell
 * at once, which is what the reference's recursion adds up to.  A task on a
 * split sub-cell (depth > 0, after scheduler_splittasks) has no counterpart
 * and the SWIFT-side shim refuses it.
 *
 * Completion: calls enqueue their work on the engine's stream and return;
 * work of later calls on the same engine runs after it (the stream carries
 * the dependencies of the task graph between device tasks).  Before host
 * code reads a task's results -- the point where SWIFT's scheduler_done
 * releases dependants that read struct part (src/scheduler.c:1203-1222) --
 * swsph_cell_write_back waits for the engine and writes the results into
 * the caller's particle memory. */

/* hydro_init_part on the active particles of the cell (called from the
 * drift/init in SWIFT, src/hydro/SPHENIX/hydro.h:564). */
int swsph_init_cell(swsph_engine *e, int ci);
/* runner_dosub_self1_density / DOSELF1_BRANCH (src/runner_doiact_hydro.h:
 * 176-177; src/runner_main.c:222). */
int swsph_doself1_density(swsph_engine *e, int ci);
/* runner_dosub_pair1_density / DOPAIR1_BRANCH (src/runner_doiact_hydro.h:
 * 181-182; src/runner_main.c:281): both directions. */
int swsph_dopair1_density(swsph_engine *e, int ci, int cj);
/* runner_doself_subset_branch_density(r, ci, parts, ind, count) and
 * runner_dopair_subset_branch_density(r, ci, parts_i, ind, count, cj)
 * (src/runner_doiact_hydro.h:197-210): the density of `count` particles of
 * cell ci against ci (resp. cj) only.  As in the reference, parts is the
 * cell's particle array inside the caller's array and ind indexes into
 * it: particle k of the subset is the caller's particle first + ind[k]
 * (first = ci->hydro.parts - s->parts).  The particles must belong to ci
 * (status 1 otherwise, nothing is run); inactive ones are skipped.  Waits
 * for the index upload and the cell check before returning. */
int swsph_doself_subset_density(swsph_engine *e, int ci, int64_t first,
                                const int *ind, int count);
int swsph_dopair_subset_density(swsph_engine *e, int ci, int64_t first,
                                const int *ind, int count, int cj);
/* runner_do_ghost(r, c, offset, ntasks, timer) (src/runner.h:94,
 * src/runner_ghost.c:1113): iterates h to convergence with subset re-runs
 * over the cell's 27 density interactions, for the share `offset` of
 * `ntasks` ghost tasks of the cell (the reference takes particles
 * offset, offset + ntasks, ... of the cell; here every ntasks-th particle
 * of the cell's device order from offset on -- each particle in exactly
 * one share).  Waits for its own convergence test, like the reference. */
int swsph_do_ghost(swsph_engine *e, int ci, int offset, int ntasks);
/* runner_dosub_self1_gradient / runner_dosub_pair1_gradient (gradient
 * instance of src/runner_doiact_hydro.h:176-185; src/runner_main.c:228,
 * 287). */
int swsph_doself1_gradient(swsph_engine *e, int ci);
int swsph_dopair1_gradient(swsph_engine *e, int ci, int cj);
/* runner_do_extra_ghost (src/runner.h:96, src/runner_ghost.c:1016). */
int swsph_do_extra_ghost(swsph_engine *e, int ci);
/* runner_dosub_self2_force / runner_dosub_pair2_force (src/runner_doiact_-
 * hydro.h:187-195; src/runner_main.c:232, 291). */
int swsph_doself2_force(swsph_engine *e, int ci);
int swsph_dopair2_force(swsph_engine *e, int ci, int cj);
/* runner_do_end_hydro_force (src/runner.h:110, src/runner_others.c:815). */
int swsph_do_end_hydro_force(swsph_engine *e, int ci);

/* A destination in the caller's memory for one field: the value of the
 * caller's particle p goes to (char *)base + p * stride, its components
 * (vectors) at consecutive elements from there -- an array of structs such
 * as struct part (base = &parts[0].rho, stride = sizeof(struct part)) or a
 * plain array (stride = the element size times the components).  Element
 * types as for swsph_get_field. */
typedef struct swsph_field_dst {
  int field;          /* enum swsph_field */
  void *base;
  int64_t stride;     /* bytes; |stride| >= the field's element bytes */
} swsph_field_dst;
/* Waits for all work enqueued on the engine, then writes the listed fields
 * of the particles of cell ci (-1: of all particles) into the caller's
 * memory: the results are in the caller's particles when it returns, as
 * SWIFT's are in struct part when a task ends (src/scheduler.c:1203-1222).
 * No counterpart in the reference, whose tasks write struct part in
 * place. */
int swsph_cell_write_back(swsph_engine *e, int ci,
                          const swsph_field_dst *dst, int nfields);

/* ---- whole passes -------------------------------------------------------
 * Plain: the C task loop of the reference's dependency order (all density
 * tasks, ghosts, gradient tasks, extra ghosts, force tasks, end-force),
 * one launch per task.  Batched: the same pass as a handful of launches over
 * all cells.  Both leave identical neighbour sets. */
int swsph_pass_plain(swsph_engine *e, swsph_pass_stats *stats);
int swsph_pass_batched(swsph_engine *e, swsph_pass_stats *stats);
/* The batched pass replayed from two captured HIP graphs: hydro_init_part,
 * the first density round and as many ghost / density re-run rounds as the
 * previous pass needed (rounds with nothing left to redo run empty), then
 * gradient -> extra ghost -> force -> end force.  The host reads one
 * convergence counter between the graphs and continues eagerly if more
 * rounds are needed; graphs are re-captured after any change of particles
 * or settings.  Results are identical to swsph_pass_batched.  Needs the
 * neighbour lists disabled.  No counterpart in the reference. */
int swsph_pass_graph(swsph_engine *e, swsph_pass_stats *stats);
/* swsph_pass_batched without any device allocation: the pass's scratch
 * (the work queue: particle flags, scans, compact particle list, tile
 * table and queue heads; the ghost and interaction counters) lives in one
 * fixed device array of the library (about 30 MB per GPU, held from the
 * first engine creation on that GPU until the process ends) instead of the
 * engine's buffers.  Same kernels and launches, so the results are
 * bit-identical to swsph_pass_batched.  Limits: at most 2^21 particles on
 * at most 2^16 top-level cells (status 1 beyond them); neighbour lists and
 * kernel timing must be off (status 1); a chunk table that would need
 * device memory (first pass after swsph_set_particles grew the set, or a
 * new tile shape) is status 1 -- run swsph_pass_batched once first.  All
 * calls on one GPU share the array: calls from different engines that
 * overlap in time on the same GPU race and must be serialised by the
 * caller.  No counterpart in the reference: for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc). */
int swsph_pass_batched_noalloc(swsph_engine *e, swsph_pass_stats *stats);

/* Kernel support larger than a top-level cell.  The whole-pass forms
 * (swsph_pass_batched, swsph_pass_graph, swsph_sched_pass,
 * swsph_engine_step) then regrid as space_regrid does
 * (src/space_regrid.c:118-126): cdim = floor(dim / (H_max * 1.10)) per
 * axis (space_stretch, src/space.h:74), never finer than before, the
 * particles re-binned on the device and the pass repeated from the
 * smoothing lengths found so far; fewer than 3 cells per axis is status 4.
 * Cell indices then refer to the new grid (swsph_num_cells).  The per-task
 * entry points, task bundles and the allocation-free pass do not regrid:
 * their caller owns the cell indices, and they return status 4.  on = 0
 * turns regridding off (status 4 everywhere); default on. */
int swsph_set_regrid(swsph_engine *e, int on);
int swsph_get_grid(swsph_engine *e, int ncell[3]);

/* Device-to-device save / restore of the state a pass reads and rewrites
 * (h and the (io) fields), on the engine stream: the device-resident
 * equivalent of restart_write_blocks / restart_read for this path
 * (src/restart.c:130, 217). */
int swsph_save_state(swsph_engine *e);
int swsph_restore_state(swsph_engine *e);

/* Tile shape of a neighbour loop (which = 0 density, 1 gradient, 2 force):
 * particles are grouped in chunks of C consecutive particles of a cell
 * (C in {8, 16, 32, 64}); one wave64 owns one chunk of i-particles with
 * 64 / C candidate streams per particle.  No counterpart in the reference
 * (its loops are scalar); the shape changes speed, not results beyond FP32
 * summation order. */
int swsph_set_chunk(swsph_engine *e, int which, int C);
int swsph_get_chunk(swsph_engine *e, int which, int *C);
/* The same with the candidate chunks of the work-queue schedules (modes 1
 * and 2 of swsph_set_schedule) sized apart from the tiles: tiles of CI
 * i-particles against candidate chunks of CJ particles, CJ = CI or
 * CJ = CI / 2 for CI = 16, 32 (finer candidate culling under the same tile),
 * the narrow tiles (4, 16), (4, 32), (8, 16) for steps with few active
 * particles per cell (64 / CI candidate streams per particle), or (16, 32)
 * (coarser chunks: fewer chunk visits, looser culling).
 * swsph_set_chunk(e, which, C) is swsph_set_tile_shape(e, which, C, C). */
int swsph_set_tile_shape(swsph_engine *e, int which, int CI, int CJ);

/* A task released by the scheduler: a self task on cell ci, or a pair task
 * on the adjacent cells ci, cj (both directions, as runner_dopair1_* /
 * runner_dopair2_*). */
enum { SWSPH_TASK_SELF = 0, SWSPH_TASK_PAIR = 1 };
typedef struct swsph_task {
  int type;
  int ci, cj;
} swsph_task;
/* Runs a bundle of released tasks of one loop (0 density, 1 gradient,
 * 2 force) in one launch: the scheduler -> device task export of a
 * task-based engine (scheduler_done releasing tasks, src/scheduler.c:1203;
 * the hydro loop tasks of engine_make_hydroloop_tasks_mapper,
 * src/engine_maketasks.c:3501).  The same interactions as
 * swsph_doself1_density / swsph_dopair1_density ... for each task, merged per
 * i-cell (one wave per chunk of every cell the bundle touches); a cell pair
 * given twice is an error.  Bundles run in call order on the engine's
 * stream, which carries the dependencies between them (density bundle ->
 * swsph_do_ghost -> gradient bundle -> ...).  Status 0 on success. */
int swsph_run_tasks(swsph_engine *e, int loop, const swsph_task *tasks,
                    int ntasks);
/* Bundles of cell tasks: runner_do_ghost, runner_do_extra_ghost,
 * runner_do_end_hydro_force and the hydro_init_part of the drift/init
 * (src/runner_ghost.c:1113, 1016; src/runner_others.c:815;
 * src/hydro/SPHENIX/hydro.h:564) on each of the n listed cells, released
 * together by the scheduler: one set of device launches over the cells'
 * particles (the ghost in device-wide rounds with re-runs over each
 * particle's 27-cell stencil).  *rounds (if not null) receives the largest
 * number of ghost rounds so far.  Status 3 if a smoothing length does not
 * converge, as swsph_do_ghost. */
int swsph_init_cells(swsph_engine *e, const int *cells, int n);
int swsph_do_ghost_cells(swsph_engine *e, const int *cells, int n,
                         int *rounds);
int swsph_do_extra_ghost_cells(swsph_engine *e, const int *cells, int n);
int swsph_do_end_hydro_force_cells(swsph_engine *e, const int *cells, int n);
/* act[c] = 1 for the cells holding an active particle (cell_is_active_-
 * hydro, src/active.h), else 0; act has swsph_num_cells entries. */
int swsph_cell_active(swsph_engine *e, uint8_t *act);

/* One density -> ghost -> gradient -> extra ghost -> force -> end force
 * pass driven by the reference's task graph instead of the device work
 * queue: the hydro tasks of the cells with active particles and their
 * dependencies (engine_make_hydroloop_tasks_mapper and engine_make_hydro_-
 * loops_dependencies, src/engine_maketasks.c:3501, 2546-2573; pair tasks
 * active when either cell is, as engine_marktasks), run by releasing each
 * set of ready tasks as one bundle per task type (scheduler_done,
 * src/scheduler.c:1203-1222) through swsph_run_tasks and the bundles of
 * cell tasks above.  Same interactions as swsph_pass_batched (FP32 sums in
 * another order). */
typedef struct swsph_sched_stats {
  int64_t active_cells;  /* cells with active particles */
  int64_t tasks;         /* tasks in the graph */
  int64_t bundles;       /* bundles submitted */
  int levels;            /* release rounds */
  double wall_ms;        /* wall time of the pass (host and device) */
} swsph_sched_stats;
int swsph_sched_pass(swsph_engine *e, swsph_sched_stats *sched,
                     swsph_pass_stats *stats);
/* 1: swsph_engine_step runs its pass through swsph_sched_pass's task graph
 * (tasks streamed from the scheduler); 0 (default): through the device
 * work queue. */
int swsph_set_step_sched(swsph_engine *e, int on);

/* Tile shape of the passes of swsph_engine_step when its active particles
 * number fewer than per_cell per top-level cell on average (work-queue
 * schedules only): all three loops then use tiles of CI against chunks of
 * CJ (default (4, 8) below 16 per cell; per_cell = 0 turns it off).  No
 * counterpart in the reference; speed only, as swsph_set_tile_shape. */
int swsph_set_sparse_shape(swsph_engine *e, int CI, int CJ, int per_cell);

/* In-cell particle order: 0 Morton order inside each top-level cell; 1
 * (default) the fields by time bin first (shortest steps first), then Morton
 * order, so that the particles active at any step form a prefix of every
 * cell and the per-particle kernels touch contiguous runs of their fields,
 * while the loops' candidate records and chunk tables keep the Morton order
 * of each cell (through a permutation pair set at every binning; none when
 * all particles share one time bin).  Resident particles are re-binned at
 * once (saved states and captured graphs are dropped); later binnings keep
 * the order, which time-step changes between binnings only loosen.  No
 * counterpart in the reference; speed only (the results differ by FP32
 * summation order at most). */
int swsph_set_bin_order(swsph_engine *e, int on);
int swsph_get_tile_shape(swsph_engine *e, int which, int *CI, int *CJ);

/* Schedule of the batched pass's neighbour loops.  mode 0: one wave per
 * chunk of every cell, launched as one grid (inactive chunks exit at once).
 * Modes 1 and 2: the particles a loop processes (all active ones, or the
 * ones the ghost sent back) are compacted on the device into a work queue
 * of tiles of <= C particles of one cell; mode 1 (default) runs one wave per
 * tile, mode 2 a persistent grid of `waves` waves (0: enough to fill every
 * CU) that drains the queue through per-XCD sub-queues with stealing.  The
 * batched form's answer to the reference's task queue (src/scheduler.c)
 * over cells of very uneven occupancy; results do not depend on the mode. */
int swsph_set_schedule(swsph_engine *e, int mode, int waves);

/* Autotuner of the batched pass: times every schedule (0, 1, 2 of
 * swsph_set_schedule) with chunk capacities 8, 16 and 32 on the engine's
 * current particles (`repeats` timed passes each, HIP events per loop),
 * keeps the schedule with the smallest summed loop time and, under it, the
 * fastest capacity of each loop; under a work-queue schedule it then tries
 * candidate chunks of half the tile size per loop, and applies the result.
 * The pass state (h and the (io) fields) is saved before and restored
 * after, so results are those of an untuned engine.  No counterpart in the
 * reference. */
typedef struct swsph_tune_result {
  int schedule;       /* chosen schedule */
  int chunk[3];       /* chosen chunk capacity of each loop */
  double loop_ms[3];  /* loop times per pass of the choice (ms) */
  double pass_ms;     /* wall time per pass, chosen schedule, capacity 16 */
  int trials;         /* configurations timed */
  int chunk_j[3];     /* chosen candidate chunk capacity of each loop */
} swsph_tune_result;
int swsph_autotune(swsph_engine *e, int repeats, swsph_tune_result *out);

/* Neighbour lists of the batched pass.  With lists enabled, the first
 * density round keeps, per particle, the candidates within `skin` times its
 * search radius (`slots` entries per candidate stream, 4 streams); the
 * ghost's density re-runs and the gradient loop then walk those lists
 * instead of the 27-cell stencil, and a particle whose radius outgrows its
 * list (or whose list overflowed) goes through the stencil loop again.
 * The role of SWIFT's cell sorts and sub-cell recursion
 * (src/runner_sort.c, src/runner_doiact_functions_hydro.h) in pruning
 * candidates; neighbour sets and results do not depend on the setting.
 * Defaults: disabled, skin 1.2, 48 slots. */
int swsph_set_lists(swsph_engine *e, int enable, float skin, int slots);

/* ---- time integration ---------------------------------------------------
 * Gas-only, non-cosmological time integration on SWIFT's integer time line
 * (src/timeline.h: a particle in bin b takes steps of 2^(b+1) time-line
 * units of props.time_base; it is active when the current time is a
 * multiple of its step).  Particles keep their state in HBM from step to
 * step; the positions are re-binned on the device whenever a drifted
 * particle could have left the 27-cell stencil of its cell (kernel support
 * H plus twice the largest distance of a particle outside its cell above
 * the cell width, the reference's rebuild criterion). */
typedef struct swsph_integration {
  float cfl;               /* SPH:CFL_condition */
  float log_max_h_change;  /* log of the largest h change per step
                              (hydro_props.log_max_h_change) */
  float min_u;             /* hydro_props.minimal_internal_energy */
  double dt_min, dt_max;   /* TimeIntegration:dt_min / dt_max */
  int fixed_bins;          /* 1: keep the time bins the particles were given
                              (benchmark hierarchies), 0: SWIFT's time-step
                              rule */
} swsph_integration;
int swsph_set_integration(swsph_engine *e, const swsph_integration *p);

/* runner_do_drift_part (src/runner_drift.c; drift_part, src/drift.h:138)
 * for cell ci (-1: all cells) from ti_old to ti_current. */
int swsph_do_drift_part(swsph_engine *e, int ci, int64_t ti_old,
                        int64_t ti_current);
/* runner_do_kick1 / runner_do_kick2 (src/runner_time_integration.c:87,
 * 360) for cell ci (-1: all cells) at ti_current. */
int swsph_do_kick1(swsph_engine *e, int ci, int64_t ti_current);
int swsph_do_kick2(swsph_engine *e, int ci, int64_t ti_current);
/* runner_do_timestep (src/runner_time_integration.c:639) for cell ci (-1:
 * all cells): new bins of the active particles; *ti_end_min receives the
 * next synchronisation point over the cell. */
int swsph_do_timestep(swsph_engine *e, int ci, int64_t ti_current,
                      int64_t *ti_end_min);
/* Re-bins the particles on the device (the particle sort of
 * space_rebuild, src/space.c, for this path). */
int swsph_rebuild(swsph_engine *e);

typedef struct swsph_step_info {
  int64_t ti_current;      /* time of the step just taken */
  int64_t ti_end_min;      /* next synchronisation point */
  int max_active_bin;
  int rebuilt;             /* particles were re-binned this step */
  int64_t n_active;
  float dx_max;            /* largest distance outside a cell after the drift */
  int64_t launches;
  swsph_pass_stats pass;
  int sparse;              /* the pass ran with the sparse tile shape */
} swsph_step_info;
/* One step of the gas (engine_step, src/engine.c, with the hydro tasks
 * only): the first call computes forces at t = 0 for all particles, their
 * time bins and the first half-kick; every later call advances to the next
 * synchronisation point, drifts all particles, re-bins if needed, runs the
 * batched density -> ghost -> gradient -> force pass over the active
 * particles, then kick2 -> time step -> kick1. */
int swsph_engine_step(swsph_engine *e, swsph_step_info *info);

/* ---- serving -------------------------------------------------------------
 * A request server for independent periodic boxes (no counterpart in the
 * reference, whose engine steps one box).  Each request is one batched
 * pass (density -> ghost -> gradient -> force) of one box.  Requests are
 * admitted first come, first served while a slot (of max_inflight) is free
 * and the device bytes of the admitted boxes stay within hbm_budget_bytes
 * (<= 0: 80 % of the free HBM at creation); every admitted request runs on
 * its own engine and HIP stream, so the work of concurrent requests
 * overlaps on the GPU, and a finished request frees its slot for the next
 * at once.  Each slot keeps its engine between requests and reuses its
 * device buffers when the next box fits (the budget counts the admitted
 * requests, not this retained capacity).  A box submitted with keep != 0 stays resident after its step
 * (its id is the request's id) and is stepped again with
 * swsph_server_step without a new upload.  Results stay on the host until
 * swsph_server_release.  Errors: swsph_server_last_error(). */
typedef struct swsph_server swsph_server;
typedef struct swsph_server_opts {
  double hbm_budget_bytes;
  int max_inflight;     /* concurrent requests (1..64) */
  int use_graphs;       /* replay each pass from captured HIP graphs */
  unsigned int fields;  /* bit f: return swsph_field f (0: all) */
} swsph_server_opts;
typedef struct swsph_request_info {
  int64_t box;           /* resident box id (0: not kept) */
  int ghost_rounds;
  int64_t n_density, n_force;
  double t_submit_ms, t_admit_ms, t_start_ms, t_done_ms;  /* steady clock */
  double t_uploaded_ms;  /* engine ready and particles on the device */
  double t_passed_ms;    /* pass finished, before the results are copied */
} swsph_request_info;
const char *swsph_server_last_error(void);
int swsph_server_create(const swsph_server_opts *opts, swsph_server **out);
/* Finishes the admitted requests; requests still queued fail. */
int swsph_server_destroy(swsph_server *s);
/* Copies the particles and queues one step; returns at once. */
int swsph_server_submit(swsph_server *s, const swsph_props *props,
                        const swsph_host_parts *parts, int keep, int64_t *id);
int swsph_server_step(swsph_server *s, int64_t box, int64_t *id);
/* Blocks until the request is done; status 3 if it failed. */
int swsph_server_wait(swsph_server *s, int64_t id, swsph_request_info *info);
/* One result field (as swsph_get_field) of a finished request. */
int swsph_server_get(swsph_server *s, int64_t id, int field, void *host_out);
int swsph_server_release(swsph_server *s, int64_t id);
int swsph_server_close_box(swsph_server *s, int64_t box);
int swsph_server_stats(swsph_server *s, double *bytes_in_use, int *inflight,
                       int *queued);

/* Interaction counters summed over active particles (device reduction). */
int swsph_count_interactions(swsph_engine *e, int64_t out3[3]);

/* The HIP stream (a hipStream_t) every launch of the engine goes to. */
int swsph_stream(swsph_engine *e, void **stream);
/* Kernel timing with HIP events recorded on the engine stream around every
 * neighbour-loop launch of the batched pass.  swsph_get_timing returns, per
 * loop (0 density incl. ghost re-runs, 1 gradient, 2 force), the summed
 * kernel time in ms and the number of launches since timing was enabled. */
int swsph_set_timing(swsph_engine *e, int on);
int swsph_get_timing(swsph_engine *e, double ms[3], int64_t launches[3]);

#ifdef __cplusplus
}
#endif
#endif
