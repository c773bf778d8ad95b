// sph_sched.h -- the hydro task graph of one step and its release to the
// device as bundles (internal interface of sph_sched.cpp).
#pragma once

#include <vector>

#include "../../include/swsph.h"

namespace sphs {

// The device work a scheduler releases, as bundles: done by the engine
// (every call enqueues on the engine's stream, which orders the bundles).
struct DeviceTasks {
  virtual ~DeviceTasks() = default;
  // act[c] != 0 for the cells holding an active particle (ncells entries)
  virtual int cell_active(std::vector<unsigned char> &act) = 0;
  virtual int init_cells(const std::vector<int> &cells) = 0;
  // loop 0 density, 1 gradient, 2 force
  virtual int loop_tasks(int loop, const std::vector<swsph_task> &tasks) = 0;
  virtual int ghost_cells(const std::vector<int> &cells, int *rounds) = 0;
  virtual int extra_ghost_cells(const std::vector<int> &cells) = 0;
  virtual int end_force_cells(const std::vector<int> &cells) = 0;
};

// Builds the hydro task graph of the cells with active particles on a
// periodic grid of ncell[3] top-level cells and runs it through `dev`,
// releasing each round of ready tasks as one bundle per task type.
int run_hydro_graph(DeviceTasks &dev, const int ncell[3],
                    swsph_sched_stats *stats, int *ghost_rounds);

}  // namespace sphs
