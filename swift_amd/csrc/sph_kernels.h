// sph_kernels.h -- device data layout and launchers shared by the engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Neighbour records: the fields a neighbour loop reads from particle j,
// packed so that the interaction phase fetches them with a few 16-byte
// loads.  The leading double position (and, for the force loop, the search
// radius) is what the candidate test reads.  Inhibited particles get a far
// position (and a negative radius) so that every test rejects them.
struct __attribute__((aligned(16))) DensRec {
  double x, y, z;
  float vx, vy, vz, m;
  float pad[2];
};
// Record sizes are odd multiples of 16 bytes so that lanes gathering
// different records from LDS spread over the banks.
struct __attribute__((aligned(16))) GradRec {
  double x, y, z;
  float vx, vy, vz, mrho;  // m_j / rho_j
  float u, c, alpha, pad0;
  float pad[6];
};
struct __attribute__((aligned(16))) ForceRec {
  double x, y, z;
  float hg2;               // (gamma_K h_j)^2
  float vx, vy, vz;
  float m, inv_m, h_inv, h_inv4;
  float rho, inv_rho, pr2, f;  // pr2 = P / rho^2
  float bal, c, alpha, P;
  float Pad, u, tb, pad0;      // Pad = P * alpha_diff
  float pad[2];
};

// Position of a candidate for the pair test, and its search radius
// (gamma_K h)^2 (read by the force test).
struct __attribute__((aligned(16))) PosRec {
  double x, y, z;
  float hg2, pad;
};

// Chunks: runs of <= C consecutive particles of one cell (Morton order inside
// the cell), with their bounding box and largest search radius.
struct ChunkArgs {
  const int *chunk_start;   // [nchunks + 1] first particle of each chunk
  const int *cell_chunk;    // [ncells + 1] first chunk of each cell
  const int *chunk_cell;    // [nchunks] cell of each chunk
  const double *bbox;       // [6 nchunks] lo x,y,z, hi x,y,z
  const float *hg2max;      // [nchunks] max (gamma_K h)^2
  int nchunks;
  int C;                    // chunk capacity (8, 16, 32 or 64)
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
};

struct GridDesc {
  int nc[3];
  double dim[3];
};

struct LoopArgs {
  GridDesc grid;
  const int *cell_start;  // [ncells + 1]
  int max_active_bin;
};

// Up to two directed cell interactions (i-cell gathers from j-cell, the
// image of the j-cell being translated by t).
struct TaskArgs {
  int icell[2];
  int jcell[2];
  int ntiles[2];
  double t[6];
};

struct GhostArgs {
  float eta_dim, eps, h_max, h_min;
  int max_active_bin;
  double time_base;
  float alpha_max, alpha_min, visc_length;
  float diff_beta, diff_alpha_max, diff_alpha_min;
};

enum class LoopKind { Density, Gradient, Force };

namespace sphk {
hipError_t launch_stencil(LoopKind kind, bool subset, const DevParts &d,
                          const LoopArgs &la, const ChunkArgs &ca,
                          hipStream_t st);
hipError_t launch_task(LoopKind kind, bool subset, const DevParts &d,
                       const LoopArgs &la, const ChunkArgs &ca,
                       const TaskArgs &ta, hipStream_t st);
// Positions / search radii of the candidate test for particles [lo, hi).
hipError_t launch_build_pos(const DevParts &d, int lo, int hi, hipStream_t st);
hipError_t launch_zero_u32(unsigned int *p, int n, hipStream_t st);
// Bounding boxes and largest search radius of chunks [c0, c1); the largest
// search radius over them is max-reduced into *hg2_all (float bits).
hipError_t launch_chunk_bounds(const DevParts &d, const int *chunk_start,
                               int c0, int c1, double *bbox, float *hg2max,
                               unsigned int *hg2_all, hipStream_t st);
hipError_t launch_init_density(const DevParts &d, int lo, int hi,
                               int max_active_bin, float h_max, hipStream_t st);
hipError_t launch_ghost(const DevParts &d, const GhostArgs &ga, int lo, int hi,
                        int *redo_count, hipStream_t st);
hipError_t launch_extra_ghost(const DevParts &d, const GhostArgs &ga, int lo,
                              int hi, hipStream_t st);
hipError_t launch_end_force(const DevParts &d, int lo, int hi,
                            int max_active_bin, hipStream_t st);
hipError_t launch_count(const DevParts &d, int n, int max_active_bin,
                        unsigned long long *out3, hipStream_t st);
hipError_t launch_build_records(LoopKind kind, const DevParts &d, int lo,
                                int hi, hipStream_t st);
}  // namespace sphk
