// sph_sched.cpp -- the hydro task graph of one step, released to the device
// as bundles: the scheduler -> device task export of a task-based engine.
//
// The graph is the reference's for the hydro loops on top-level cells, for
// the cells with active particles (engine_marktasks activates a pair task
// when either of its cells is active):
//   per active cell c: init(c) (hydro_init_part from the drift/init),
//     ghost(c), extra_ghost(c), end_force(c);
//   per loop (density, gradient, force): a self task on every active cell
//     and a pair task on every pair of adjacent cells with an active side
//     (engine_make_hydroloop_tasks_mapper, src/engine_maketasks.c:3501);
//   dependencies (engine_make_hydro_loops_dependencies,
//     src/engine_maketasks.c:2546-2573, and the init / end-force links):
//     init(c) -> density tasks of c -> ghost(c) -> gradient tasks of c
//     -> extra_ghost(c) -> force tasks of c -> end_force(c),
//     a pair task waiting for the cell tasks of both active cells and
//     unlocking both.
// Execution: every task carries the count of its unreleased dependencies;
// the tasks at zero form the ready set, which is submitted as one bundle
// per task type (loop tasks through swsph_run_tasks, cell tasks through the
// multi-cell entry points), after which the tasks they unlock are released
// (scheduler_done, src/scheduler.c:1203-1222).  A bundle only enqueues work
// on the engine's stream, whose order keeps every task behind the work it
// depends on, so dependants are released as soon as their dependencies are
// enqueued.
//
// The graph is regular, so its edges are implicit: loop task (l, c, s) is
// the self task of cell c (s = 0) or the pair task of c and its neighbour
// at stencil offset 13 + s (s = 1..13, the "upper" half, so that every
// adjacent pair has one task); cell task (k, c) is init / ghost / extra
// ghost / end force of c.  Only the dependency counters are stored.
#include "sph_sched.h"

#include <chrono>
#include <cstdint>

namespace sphs {
namespace {

constexpr int kSlots = 14;  // self + 13 upper pairs per cell and loop

struct Grid {
  int n[3];
  int cells() const { return n[0] * n[1] * n[2]; }
  // neighbour of c at stencil offset o = 9 (ox+1) + 3 (oy+1) + (oz+1)
  int nb(int c, int o) const {
    const int cx = c / (n[1] * n[2]), cy = (c / n[2]) % n[1], cz = c % n[2];
    const int x = (cx + o / 9 - 1 + n[0]) % n[0];
    const int y = (cy + (o / 3) % 3 - 1 + n[1]) % n[1];
    const int z = (cz + o % 3 - 1 + n[2]) % n[2];
    return (x * n[1] + y) * n[2] + z;
  }
};

}  // namespace

int run_hydro_graph(DeviceTasks &dev, const int ncell[3],
                    swsph_sched_stats *stats, int *ghost_rounds) {
  const auto t0 = std::chrono::steady_clock::now();
  const Grid g{{ncell[0], ncell[1], ncell[2]}};
  const int nc = g.cells();
  std::vector<unsigned char> act;
  int rc = dev.cell_active(act);
  if (rc) return rc;
  if ((int)act.size() != nc) return 2;
  // Dependency counters: loop tasks [l][c * kSlots + s] (-1: no such task),
  // cell tasks [k][c] (k: 0 init, 1 ghost, 2 extra ghost, 3 end force).
  std::vector<int> lwait[3], cwait[4];
  for (auto &v : lwait) v.assign((size_t)nc * kSlots, -1);
  for (auto &v : cwait) v.assign((size_t)nc, -1);
  int64_t ntasks = 0, nactive = 0;
  for (int c = 0; c < nc; c++) {
    if (act[c]) {
      nactive++;
      cwait[0][c] = 0;
      // ghost / extra ghost / end force wait for the 27 loop tasks of c
      for (int k = 1; k < 4; k++) cwait[k][c] = 27;
      ntasks += 4;
    }
    for (int s = 0; s < kSlots; s++) {
      const int d = s == 0 ? c : g.nb(c, 13 + s);
      const int w = (act[c] ? 1 : 0) + (s > 0 && act[d] ? 1 : 0);
      if (w == 0) continue;
      for (int l = 0; l < 3; l++) lwait[l][(size_t)c * kSlots + s] = w;
      ntasks += 3;
    }
  }
  // Ready sets: cell tasks of kind k and loop tasks of loop l.
  std::vector<int> rcell[4], rloop[3], ncell_r[4], nloop_r[3];
  for (int c = 0; c < nc; c++)
    if (act[c]) rcell[0].push_back(c);
  // Releases the loop tasks of loop l that touch cell c (a cell task of c
  // before them completed).
  auto release_loops = [&](int l, int c, std::vector<int> &out) {
    auto dec = [&](size_t id) {
      if (lwait[l][id] > 0 && --lwait[l][id] == 0) out.push_back((int)id);
    };
    dec((size_t)c * kSlots);
    for (int o = 14; o < 27; o++) {
      dec((size_t)c * kSlots + (o - 13));          // (c, upper neighbour)
      dec((size_t)g.nb(c, 26 - o) * kSlots + (o - 13));  // (lower, c)
    }
  };
  // Releases the cell task of kind k of cell c (a loop task touching it
  // completed).
  auto release_cell = [&](int k, int c, std::vector<int> &out) {
    if (cwait[k][c] > 0 && --cwait[k][c] == 0) out.push_back(c);
  };
  int64_t bundles = 0, released = 0;
  int levels = 0;
  std::vector<swsph_task> bundle;
  for (;;) {
    bool any = false;
    for (auto &v : rcell) any = any || !v.empty();
    for (auto &v : rloop) any = any || !v.empty();
    if (!any) break;
    levels++;
    for (auto &v : ncell_r) v.clear();
    for (auto &v : nloop_r) v.clear();
    // cell tasks: init -> density tasks, ghost -> gradient tasks, extra
    // ghost -> force tasks, end force -> nothing
    if (!rcell[0].empty()) {
      if ((rc = dev.init_cells(rcell[0]))) return rc;
      bundles++;
    }
    if (!rcell[1].empty()) {
      if ((rc = dev.ghost_cells(rcell[1], ghost_rounds))) return rc;
      bundles++;
    }
    if (!rcell[2].empty()) {
      if ((rc = dev.extra_ghost_cells(rcell[2]))) return rc;
      bundles++;
    }
    if (!rcell[3].empty()) {
      if ((rc = dev.end_force_cells(rcell[3]))) return rc;
      bundles++;
    }
    for (int l = 0; l < 3; l++) {
      if (rloop[l].empty()) continue;
      bundle.clear();
      bundle.reserve(rloop[l].size());
      for (int id : rloop[l]) {
        const int c = id / kSlots, s = id % kSlots;
        swsph_task t;
        t.type = s == 0 ? SWSPH_TASK_SELF : SWSPH_TASK_PAIR;
        t.ci = c;
        t.cj = s == 0 ? c : g.nb(c, 13 + s);
        bundle.push_back(t);
      }
      if ((rc = dev.loop_tasks(l, bundle))) return rc;
      bundles++;
    }
    // scheduler_done: release the dependants
    for (int k = 0; k < 3; k++)
      for (int c : rcell[k]) release_loops(k, c, nloop_r[k]);
    for (int l = 0; l < 3; l++)
      for (int id : rloop[l]) {
        const int c = id / kSlots, s = id % kSlots;
        if (act[c]) release_cell(l + 1, c, ncell_r[l + 1]);
        if (s > 0) {
          const int d = g.nb(c, 13 + s);
          if (act[d]) release_cell(l + 1, d, ncell_r[l + 1]);
        }
      }
    for (int k = 0; k < 4; k++) released += (int64_t)rcell[k].size();
    for (int l = 0; l < 3; l++) released += (int64_t)rloop[l].size();
    for (int k = 0; k < 4; k++) rcell[k].swap(ncell_r[k]);
    for (int l = 0; l < 3; l++) rloop[l].swap(nloop_r[l]);
  }
  if (stats) {
    stats->active_cells = nactive;
    stats->tasks = ntasks;
    stats->bundles = bundles;
    stats->levels = levels;
    stats->wall_ms = std::chrono::duration<double, std::milli>(
                         std::chrono::steady_clock::now() - t0)
                         .count();
  }
  // every task released (no dependency left waiting)
  return released == ntasks ? 0 : 2;
}

}  // namespace sphs
