// sph_kernels.hip -- gfx950 kernels of the SPHENIX neighbour pass.
//
// Data layout (HBM, structure of arrays, particles sorted by top-level cell):
//   px,py,pz     double  positions (only used to form pair offsets)
//   vx,vy,vz,m,u,h       float
//   tb                   int8 time bin
//   per-loop accumulators and prepared force quantities: float arrays
//
// One wave64 owns a "tile" = up to 64 consecutive particles of one cell, one
// particle per lane.  A loop kernel walks a list of neighbour cells; each
// cell is streamed through LDS in chunks of 64 particles with one coalesced
// load per lane, the j record being converted once at staging (e.g. m/rho,
// P/rho^2, h^-4).  Every lane then tests the chunk against its own particle
// (broadcast LDS reads) and accumulates in registers.  Accumulators are read
// from HBM at the start and written back at the end, so the same kernel
// serves the per-task drop-in (one neighbour cell per launch, the running
// sums live in HBM between tasks) and the batched pass (27 neighbour cells in
// one launch).  No atomics: every directed interaction is evaluated by the
// lane that owns particle i (gather form of DOPAIR1/DOSELF1/DOPAIR2/DOSELF2,
// src/runner_doiact_functions_hydro.h:1234, 2299, 1601, 2624).
#include "sph_device.h"
#include "sph_kernels.h"

using namespace sph;

namespace {

__device__ __forceinline__ bool is_active(int8_t tb, int max_active_bin) {
  return tb <= max_active_bin && tb != kBinInhibited;  // active.h:349
}

// Positions of an inhibited neighbour are replaced by a far sentinel at
// staging so that every mask rejects it (active.h:518 part_is_inhibited).
constexpr double kFar = 1.0e30;

struct JDens {
  double x, y, z;
  float vx, vy, vz, m;
};
struct JGrad {
  double x, y, z;
  float vx, vy, vz, mrho;  // m_j / rho_j
  float u, c, alpha, pad;
};
struct JForce {
  double x, y, z;
  float vx, vy, vz, m;
  float inv_m, hg2, h_inv, h_inv4;
  float inv_rho, pr2, f, bal;
  float c, alpha, P, Pad;
  float u, tb, rho, pad;
};

// ---------------------------------------------------------------- density --
struct DensAcc {
  float rho, rho_dh, wc, wc_dh, divv, rx, ry, rz;
  int n;
};
struct DensI {
  double x, y, z;
  float vx, vy, vz, h_inv, hg2;
};

struct DensityLoop {
  using J = JDens;
  using Acc = DensAcc;
  using I = DensI;
  static __device__ void load_i(const DevParts &d, int i, I &s) {
    s.x = d.px[i];
    s.y = d.py[i];
    s.z = d.pz[i];
    s.vx = d.vx[i];
    s.vy = d.vy[i];
    s.vz = d.vz[i];
    const float h = d.h[i];
    s.h_inv = 1.f / h;
    s.hg2 = search_radius2(h);
  }
  static __device__ void load_acc(const DevParts &d, int i, Acc &a) {
    a.rho = d.rho[i];
    a.rho_dh = d.rho_dh[i];
    a.wc = d.wc[i];
    a.wc_dh = d.wc_dh[i];
    a.divv = d.divv[i];
    a.rx = d.rotx[i];
    a.ry = d.roty[i];
    a.rz = d.rotz[i];
    a.n = d.nd[i];
  }
  static __device__ void store_acc(const DevParts &d, int i, const Acc &a) {
    d.rho[i] = a.rho;
    d.rho_dh[i] = a.rho_dh;
    d.wc[i] = a.wc;
    d.wc_dh[i] = a.wc_dh;
    d.divv[i] = a.divv;
    d.rotx[i] = a.rx;
    d.roty[i] = a.ry;
    d.rotz[i] = a.rz;
    d.nd[i] = a.n;
  }
  static __device__ void stage(const DevParts &d, int j, J &s) {
    const bool inh = d.tb[j] == kBinInhibited;
    s.x = inh ? kFar : d.px[j];
    s.y = inh ? kFar : d.py[j];
    s.z = inh ? kFar : d.pz[j];
    s.vx = d.vx[j];
    s.vy = d.vy[j];
    s.vz = d.vz[j];
    s.m = d.m[j];
  }
  static __device__ __forceinline__ bool hit(float r2, const I &s, const J &) {
    return r2 < s.hg2;
  }
  // runner_iact_nonsym_density (hydro_iact.h:141-193)
  static __device__ __forceinline__ void interact(float r2, float dx, float dy,
                                                  float dz, const I &s,
                                                  const J &j, Acc &a) {
    const float r_inv = r2 > 0.f ? fast_rsq(r2) : 0.f;
    const float r = r2 * r_inv;
    const float ui = r * s.h_inv;
    float wi, wi_dx;
    kernel_deval(ui, wi, wi_dx);
    const float t = fmaf(kDim, wi, ui * wi_dx);
    a.rho = fmaf(j.m, wi, a.rho);
    a.rho_dh = fmaf(-j.m, t, a.rho_dh);
    a.wc += wi;
    a.wc_dh -= t;
    const float faci = j.m * wi_dx * r_inv;
    const float dvx = s.vx - j.vx, dvy = s.vy - j.vy, dvz = s.vz - j.vz;
    const float dvdr = fmaf(dvz, dz, fmaf(dvy, dy, dvx * dx));
    a.divv = fmaf(-faci, dvdr, a.divv);
    a.rx = fmaf(faci, fmaf(dvy, dz, -dvz * dy), a.rx);
    a.ry = fmaf(faci, fmaf(dvz, dx, -dvx * dz), a.ry);
    a.rz = fmaf(faci, fmaf(dvx, dy, -dvy * dx), a.rz);
    a.n += 1;
  }
};

// --------------------------------------------------------------- gradient --
struct GradAcc {
  float vsig, lapu, avmax;
  int n;
};
struct GradI {
  double x, y, z;
  float vx, vy, vz, h_inv, hg2, u, c;
};

struct GradientLoop {
  using J = JGrad;
  using Acc = GradAcc;
  using I = GradI;
  static __device__ void load_i(const DevParts &d, int i, I &s) {
    s.x = d.px[i];
    s.y = d.py[i];
    s.z = d.pz[i];
    s.vx = d.vx[i];
    s.vy = d.vy[i];
    s.vz = d.vz[i];
    const float h = d.h[i];
    s.h_inv = 1.f / h;
    s.hg2 = search_radius2(h);
    s.u = d.u[i];
    s.c = d.c[i];
  }
  static __device__ void load_acc(const DevParts &d, int i, Acc &a) {
    a.vsig = d.vsig[i];
    a.lapu = d.lapu[i];
    a.avmax = d.avmax[i];
    a.n = d.ng[i];
  }
  static __device__ void store_acc(const DevParts &d, int i, const Acc &a) {
    d.vsig[i] = a.vsig;
    d.lapu[i] = a.lapu;
    d.avmax[i] = a.avmax;
    d.ng[i] = a.n;
  }
  static __device__ void stage(const DevParts &d, int j, J &s) {
    const bool inh = d.tb[j] == kBinInhibited;
    s.x = inh ? kFar : d.px[j];
    s.y = inh ? kFar : d.py[j];
    s.z = inh ? kFar : d.pz[j];
    s.vx = d.vx[j];
    s.vy = d.vy[j];
    s.vz = d.vz[j];
    s.mrho = d.m[j] / d.rho[j];
    s.u = d.u[j];
    s.c = d.c[j];
    s.alpha = d.alpha[j];
  }
  static __device__ __forceinline__ bool hit(float r2, const I &s, const J &) {
    return r2 < s.hg2;
  }
  // runner_iact_nonsym_gradient (hydro_iact.h:291-344), a = 1, H = 0
  static __device__ __forceinline__ void interact(float r2, float dx, float dy,
                                                  float dz, const I &s,
                                                  const J &j, Acc &a) {
    const float r_inv = r2 > 0.f ? fast_rsq(r2) : 0.f;
    const float r = r2 * r_inv;
    const float dvdr =
        fmaf(s.vz - j.vz, dz, fmaf(s.vy - j.vy, dy, (s.vx - j.vx) * dx));
    const float mu = r_inv * fminf(dvdr, 0.f);
    a.vsig = fmaxf(a.vsig, fmaf(-kViscBeta, mu, s.c + j.c));
    const float dw = kernel_dw(r * s.h_inv);
    a.lapu = fmaf(j.mrho * (s.u - j.u) * r_inv, dw, a.lapu);
    a.avmax = fmaxf(a.avmax, j.alpha);
    a.n += 1;
  }
};

// ------------------------------------------------------------------ force --
struct ForceAcc {
  float ax, ay, az, udt, hdt;
  int n, mintb;
};
struct ForceI {
  double x, y, z;
  float vx, vy, vz, m, inv_m, hg2, h_inv, h_inv4;
  float inv_rho, pr2, f, bal, c, alpha, P, Pad, u, rho;
};

struct ForceLoop {
  using J = JForce;
  using Acc = ForceAcc;
  using I = ForceI;
  static __device__ void load_i(const DevParts &d, int i, I &s) {
    s.x = d.px[i];
    s.y = d.py[i];
    s.z = d.pz[i];
    s.vx = d.vx[i];
    s.vy = d.vy[i];
    s.vz = d.vz[i];
    s.m = d.m[i];
    s.inv_m = 1.f / s.m;
    const float h = d.h[i];
    s.hg2 = search_radius2(h);
    s.h_inv = 1.f / h;
    const float hi2 = s.h_inv * s.h_inv;
    s.h_inv4 = hi2 * hi2;
    s.rho = d.rho[i];
    s.inv_rho = 1.f / s.rho;
    s.P = d.P[i];
    s.pr2 = s.P / (s.rho * s.rho);
    s.f = d.f[i];
    s.bal = d.bal[i];
    s.c = d.c[i];
    s.alpha = d.alpha[i];
    s.Pad = s.P * d.adiff[i];
    s.u = d.u[i];
  }
  static __device__ void load_acc(const DevParts &d, int i, Acc &a) {
    a.ax = d.ax[i];
    a.ay = d.ay[i];
    a.az = d.az[i];
    a.udt = d.udt[i];
    a.hdt = d.hdt[i];
    a.n = d.nf[i];
    a.mintb = d.mintb[i];
  }
  static __device__ void store_acc(const DevParts &d, int i, const Acc &a) {
    d.ax[i] = a.ax;
    d.ay[i] = a.ay;
    d.az[i] = a.az;
    d.udt[i] = a.udt;
    d.hdt[i] = a.hdt;
    d.nf[i] = a.n;
    d.mintb[i] = (int8_t)a.mintb;
  }
  static __device__ void stage(const DevParts &d, int j, J &s) {
    const int8_t tb = d.tb[j];
    const bool inh = tb == kBinInhibited;
    s.x = inh ? kFar : d.px[j];
    s.y = inh ? kFar : d.py[j];
    s.z = inh ? kFar : d.pz[j];
    s.vx = d.vx[j];
    s.vy = d.vy[j];
    s.vz = d.vz[j];
    s.m = d.m[j];
    s.inv_m = 1.f / s.m;
    const float h = d.h[j];
    s.hg2 = inh ? -1.f : search_radius2(h);
    s.h_inv = 1.f / h;
    const float hi2 = s.h_inv * s.h_inv;
    s.h_inv4 = hi2 * hi2;
    s.rho = d.rho[j];
    s.inv_rho = 1.f / s.rho;
    s.P = d.P[j];
    s.pr2 = s.P / (s.rho * s.rho);
    s.f = d.f[j];
    s.bal = d.bal[j];
    s.c = d.c[j];
    s.alpha = d.alpha[j];
    s.Pad = s.P * d.adiff[j];
    s.u = d.u[j];
    s.tb = (float)tb;
  }
  static __device__ __forceinline__ bool hit(float r2, const I &s, const J &j) {
    return (r2 < s.hg2) || (r2 < j.hg2);
  }
  // runner_iact_nonsym_force (hydro_iact.h:507-632) + runner_iact_nonsym_
  // timebin (timestep_limiter_iact.h:59-70), a = 1, H = 0
  static __device__ __forceinline__ void interact(float r2, float dx, float dy,
                                                  float dz, const I &s,
                                                  const J &j, Acc &a) {
    const float r_inv = r2 > 0.f ? fast_rsq(r2) : 0.f;
    const float r = r2 * r_inv;
    float wi, wi_dx, wj, wj_dx;
    kernel_deval(r * s.h_inv, wi, wi_dx);
    kernel_deval(r * j.h_inv, wj, wj_dx);
    const float wi_dr = s.h_inv4 * wi_dx;
    const float wj_dr = j.h_inv4 * wj_dx;
    const float dvdr =
        fmaf(s.vz - j.vz, dz, fmaf(s.vy - j.vy, dy, (s.vx - j.vx) * dx));
    const float mu = r_inv * fminf(dvdr, 0.f);
    const float v_sig = fmaf(-kViscBeta, mu, s.c + j.c);
    const float f_ij = fmaf(-s.f, j.inv_m, 1.f);
    const float f_ji = fmaf(-j.f, s.inv_m, 1.f);
    const float rho_ij = s.rho + j.rho;
    const float inv_rho_ij = fast_rcp(rho_ij);
    const float visc = -0.25f * (s.alpha + j.alpha) * v_sig * mu *
                       (s.bal + j.bal) * inv_rho_ij;
    const float visc_acc = 0.5f * visc * fmaf(wi_dr, f_ij, wj_dr * f_ji) * r_inv;
    const float Pi = s.pr2 * f_ij;
    const float Pj = j.pr2 * f_ji;
    const float acc = fmaf(fmaf(Pi, wi_dr, Pj * wj_dr), r_inv, visc_acc);
    const float mja = j.m * acc;
    a.ax = fmaf(-mja, dx, a.ax);
    a.ay = fmaf(-mja, dy, a.ay);
    a.az = fmaf(-mja, dz, a.az);
    const float dvdr_rinv = dvdr * r_inv;
    const float sph_du = Pi * dvdr_rinv * wi_dr;
    const float visc_du = 0.5f * visc_acc * dvdr;
    const float alpha_diff = (s.Pad + j.Pad) * fast_rcp(s.P + j.P);
    const float v_diff =
        alpha_diff * 0.5f *
        (fast_sqrt(2.f * fabsf(s.P - j.P) * inv_rho_ij) + fabsf(dvdr_rinv));
    const float diff_du = v_diff * (s.u - j.u) *
                          fmaf(f_ij * wi_dr, s.inv_rho, f_ji * wj_dr * j.inv_rho);
    a.udt = fmaf(sph_du + visc_du + diff_du, j.m, a.udt);
    a.hdt = fmaf(-j.m * dvdr_rinv * j.inv_rho, wi_dr, a.hdt);
    const int tbj = (int)j.tb;
    a.mintb = (tbj > 0 && tbj < a.mintb) ? tbj : a.mintb;
    a.n += 1;
  }
};

// ------------------------------------------------------- the gather engine --
// Streams the particles [j0, j0+nj) of one neighbour cell through LDS and
// accumulates into the lane's particle.  (xs,ys,zs) = x_i - t.
template <class L>
__device__ __forceinline__ void gather_cell(const DevParts &d, int j0, int nj,
                                            double xs, double ys, double zs,
                                            const typename L::I &s, bool act,
                                            int self_index,
                                            typename L::Acc &acc,
                                            typename L::J *lds) {
  const int lane = threadIdx.x;
  for (int base = 0; base < nj; base += 64) {
    const int cnt = min(64, nj - base);
    __syncthreads();
    if (lane < cnt) L::stage(d, j0 + base + lane, lds[lane]);
    __syncthreads();
    if (act) {
      const int skip = self_index - (j0 + base);  // lane's own slot, if any
      for (int k = 0; k < cnt; k++) {
        const typename L::J &j = lds[k];
        float dx, dy, dz;
        const float r2 = pair_r2(xs, ys, zs, j.x, j.y, j.z, dx, dy, dz);
        if (L::hit(r2, s, j) && k != skip) L::interact(r2, dx, dy, dz, s, j, acc);
      }
    }
  }
}

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

// Map a block to a tile so that each XCD (blocks are dealt round-robin over
// the 8 XCDs) works on a contiguous, spatially coherent range of tiles and
// reuses neighbour cells from its own L2.
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  const int per = (ntiles + 7) >> 3;
  return (b & 7) * per + (b >> 3);
}

template <class L, bool SUBSET>
__device__ __forceinline__ void run_tile(const DevParts &d, const LoopArgs &la,
                                         int cell, int first, int count,
                                         const int *ncell_list, int nlist,
                                         const double *tlist, bool stencil) {
  __shared__ typename L::J lds[64];
  const int lane = threadIdx.x;
  const int i = first + lane;
  bool act = lane < count;
  if (act) {
    act = is_active(d.tb[i], la.max_active_bin);
    if (SUBSET) act = act && d.redo[i];
  }
  if (!__any(act)) return;  // the whole wave skips (uniform)
  typename L::I s;
  typename L::Acc acc;
  if (act) {
    L::load_i(d, i, s);
    L::load_acc(d, i, acc);
  }
  if (stencil) {
    int cx, cy, cz;
    cell_coords(cell, la.grid, cx, cy, cz);
    for (int o = 0; o < 27; o++) {
      double tx, ty, tz;
      const int cj = neighbour(la.grid, cx, cy, cz, o, tx, ty, tz);
      const int j0 = la.cell_start[cj];
      const int nj = la.cell_start[cj + 1] - j0;
      gather_cell<L>(d, j0, nj, s.x - tx, s.y - ty, s.z - tz, s, act, i, acc,
                     lds);
    }
  } else {
    for (int q = 0; q < nlist; q++) {
      const int cj = ncell_list[q];
      const int j0 = la.cell_start[cj];
      const int nj = la.cell_start[cj + 1] - j0;
      gather_cell<L>(d, j0, nj, s.x - tlist[3 * q], s.y - tlist[3 * q + 1],
                     s.z - tlist[3 * q + 2], s, act, i, acc, lds);
    }
  }
  if (act) L::store_acc(d, i, acc);
}

// Batched: every tile of the tile list against its 27-cell stencil.
template <class L, bool SUBSET>
__global__ __launch_bounds__(64) void k_loop_stencil(DevParts d, LoopArgs la,
                                                     const int4 *tiles,
                                                     int ntiles) {
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  const int4 tl = tiles[t];
  run_tile<L, SUBSET>(d, la, tl.x, tl.y, tl.z, nullptr, 0, nullptr, true);
}

// Per task: up to two directed (i-cell <- j-cell) interactions; blocks
// [0, nt0) take tiles of the first i-cell, the rest the second.
template <class L, bool SUBSET>
__global__ __launch_bounds__(64) void k_loop_task(DevParts d, LoopArgs la,
                                                  TaskArgs ta) {
  int q = 0;
  int b = blockIdx.x;
  if (b >= ta.ntiles[0]) {
    q = 1;
    b -= ta.ntiles[0];
  }
  const int ci = ta.icell[q];
  const int first = la.cell_start[ci] + 64 * b;
  const int count = min(64, la.cell_start[ci + 1] - first);
  const int cj = ta.jcell[q];
  run_tile<L, SUBSET>(d, la, ci, first, count, &cj, 1, &ta.t[3 * q], false);
}

// ------------------------------------------------------ per-particle work --
// hydro_init_part (hydro.h:564-606) for active particles in [lo, hi); also
// resets the ghost bracket.
__global__ void k_init_density(DevParts d, int lo, int hi, int max_active_bin,
                               float h_max) {
  const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  const bool act = is_active(d.tb[i], max_active_bin);
  d.redo[i] = act ? 1 : 0;
  if (!act) return;
  d.rho[i] = 0.f;
  d.rho_dh[i] = 0.f;
  d.wc[i] = 0.f;
  d.wc_dh[i] = 0.f;
  d.divv[i] = 0.f;
  d.rotx[i] = 0.f;
  d.roty[i] = 0.f;
  d.rotz[i] = 0.f;
  d.nd[i] = 0;
  d.lapu[i] = 0.f;
  d.left[i] = 0.f;
  d.right[i] = h_max;
}

// hydro_prepare_gradient + hydro_reset_gradient (hydro.h:671-750) on a
// particle whose density sums are final.
__device__ void prepare_gradient(const DevParts &d, int i, float h, float rho,
                                 float rho_dh, float wc, float wc_dh,
                                 float divv, float rx, float ry, float rz,
                                 const GhostArgs &ga) {
  const float curl_v = sqrtf(rx * rx + ry * ry + rz * rz);
  const float abs_div_v = fabsf(divv);
  const float P = kGasGm1 * d.u[i] * rho;
  const float c = sqrtf(kGasGamma * P / rho);
  const float bal = abs_div_v / (abs_div_v + curl_v + 0.0001f * c / h);
  const float common = h * kDimInv / wc;
  float fgh;
  if (h > 0.9999f * ga.h_max) {
    fgh = 0.f;
  } else {
    const float gW = common * wc_dh;
    fgh = (gW < -0.9999f) ? 0.f : common * rho_dh / (1.f + gW);
  }
  d.P[i] = P;
  d.c[i] = c;
  d.f[i] = fgh;
  d.bal[i] = bal;
  d.vsig[i] = 2.f * c;
  d.avmax[i] = d.alpha[i];
  d.lapu[i] = 0.f;
  d.ng[i] = 0;
}

// One ghost round (runner_do_ghost, runner_ghost.c:1113-1628) for the
// particles flagged in redo[] within [lo, hi): hydro_end_density, one
// Newton-Raphson / bisection step on h, and either convergence (prepare the
// gradient loop) or a re-initialised particle flagged for another density
// round.  redo_count[0] receives the number of particles flagged again.
__global__ void k_ghost(DevParts d, GhostArgs ga, int lo, int hi,
                        int *redo_count) {
  const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  if (!d.redo[i]) return;
  const float h_old = d.h[i];
  const float h_old_dim = h_old * h_old * h_old;
  const float h_old_dim_m1 = h_old * h_old;
  float rho = d.rho[i], rho_dh = d.rho_dh[i], wc = d.wc[i], wc_dh = d.wc_dh[i];
  float divv = d.divv[i], rx = d.rotx[i], ry = d.roty[i], rz = d.rotz[i];
  float left = d.left[i], right = d.right[i];
  float h_new;
  bool no_ngb = false;
  bool done = false;
  if ((double)wc < 1.e-5 * (double)kRoot) {
    no_ngb = true;
    h_new = 2.f * h_old;
  } else {
    // hydro_end_density (hydro.h:613-662), a = 1, H = 0
    const float h_inv = 1.0f / h_old;
    const float hid = h_inv * h_inv * h_inv;
    const float hid1 = hid * h_inv;
    const float mi = d.m[i];
    rho = (rho + mi * kRoot) * hid;
    rho_dh = (rho_dh - kDim * mi * kRoot) * hid1;
    wc = (wc + kRoot) * hid;
    wc_dh = (wc_dh - kDim * kRoot) * hid1;
    const float rho_inv = 1.f / rho;
    rx *= hid1 * rho_inv;
    ry *= hid1 * rho_inv;
    rz *= hid1 * rho_inv;
    divv *= hid1 * rho_inv;
    const float n_sum = wc * h_old_dim;
    const float f = n_sum - ga.eta_dim;
    const float f_prime = wc_dh * h_old_dim + kDim * wc * h_old_dim_m1;
    if (n_sum < ga.eta_dim)
      left = fmaxf(left, h_old);
    else if (n_sum > ga.eta_dim)
      right = fminf(right, h_old);
    if (((h_old >= ga.h_max) && (f < 0.f)) || ((h_old <= ga.h_min) && (f > 0.f))) {
      done = true;
      h_new = h_old;
    } else {
      h_new = h_old - f / (f_prime + 1.17549435e-38f);
      h_new = fminf(h_new, 2.f * h_old);
      h_new = fmaxf(h_new, 0.5f * h_old);
      h_new = fmaxf(h_new, left);
      h_new = fminf(h_new, right);
    }
  }
  float h_fin = h_old;
  if (!done && fabsf(h_new - h_old) > ga.eps * h_old) {
    float hn;
    if ((h_new == left && h_old == right) || (h_old == left && h_new == right))
      hn = cbrtf(0.5f * (left * left * left + right * right * right));
    else
      hn = h_new;
    if (hn < ga.h_max && hn > ga.h_min) {
      // Not converged: re-initialise (hydro_init_part) and ask for a rerun.
      d.h[i] = hn;
      d.left[i] = left;
      d.right[i] = right;
      d.rho[i] = 0.f;
      d.rho_dh[i] = 0.f;
      d.wc[i] = 0.f;
      d.wc_dh[i] = 0.f;
      d.divv[i] = 0.f;
      d.rotx[i] = 0.f;
      d.roty[i] = 0.f;
      d.rotz[i] = 0.f;
      d.nd[i] = 0;
      atomicAdd(redo_count, 1);
      return;
    }
    if (hn <= ga.h_min) {
      h_fin = ga.h_min;
    } else {
      h_fin = ga.h_max;
      if (no_ngb) {
        // hydro_part_has_no_neighbours (hydro.h:791-825)
        const float hinv = 1.0f / h_fin;
        const float hid = hinv * hinv * hinv;
        rho = d.m[i] * kRoot * hid;
        wc = kRoot * hid;
        rho_dh = 0.f;
        wc_dh = 0.f;
        rx = ry = rz = 0.f;
        divv = 0.f;
      }
    }
  }
  // Converged.
  d.h[i] = h_fin;
  d.redo[i] = 0;
  d.rho[i] = rho;
  d.rho_dh[i] = rho_dh;
  d.wc[i] = wc;
  d.wc_dh[i] = wc_dh;
  d.divv[i] = divv;
  d.rotx[i] = rx;
  d.roty[i] = ry;
  d.rotz[i] = rz;
  prepare_gradient(d, i, h_fin, rho, rho_dh, wc, wc_dh, divv, rx, ry, rz, ga);
}

// runner_do_extra_ghost (runner_ghost.c:1016-1111): hydro_end_gradient,
// hydro_prepare_force (hydro.h:840-954, a = 1), hydro_reset_acceleration and
// timestep_limiter_prepare_force for active particles in [lo, hi).
__global__ void k_extra_ghost(DevParts d, GhostArgs ga, int lo, int hi) {
  const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  const int8_t tb = d.tb[i];
  if (!is_active(tb, ga.max_active_bin)) return;
  const float h = d.h[i];
  const float h_inv = 1.0f / h;
  const float hid1 = h_inv * h_inv * h_inv * h_inv;
  const float lapu = d.lapu[i] * (2.f * hid1);
  d.lapu[i] = lapu;
  const float dt_alpha =
      tb <= 0 ? 0.f : (float)((double)(1LL << (tb + 1)) * ga.time_base);
  const float ksp = h * kGamma;
  const float ksp_inv = 1.f / ksp;
  const float vsig = d.vsig[i];
  const float rho = d.rho[i];
  const float P = kGasGm1 * d.u[i] * rho;
  const float c = sqrtf(kGasGamma * P / rho);
  const float sct_inv = c * ksp_inv;
  const float divv = d.divv[i];
  const float div_v_dt =
      dt_alpha == 0.f ? 0.f : (divv - d.divv_prev[i]) / dt_alpha;
  const float S = divv < 0.f ? ksp * ksp * fmaxf(0.f, -div_v_dt) : 0.f;
  const float alpha_loc = ga.alpha_max * S / (c * c + S);
  float alpha = d.alpha[i];
  if (alpha_loc > alpha) {
    alpha = alpha_loc;
  } else {
    const float tr = dt_alpha * sct_inv * ga.visc_length;
    alpha += alpha_loc * tr;
    alpha /= (1.f + tr);
  }
  alpha = fmaxf(alpha, ga.alpha_min);
  d.alpha[i] = alpha;
  d.divv_prev[i] = divv;
  d.divv_dt[i] = div_v_dt;
  const float dtpi = vsig * ksp_inv;
  const float sqrt_u_inv = 1.f / sqrtf(d.u[i]);
  float adt = ga.diff_beta * ksp * lapu * sqrt_u_inv;
  const float ad0 = d.adiff[i];
  adt -= (ad0 - ga.diff_alpha_min) * dtpi;
  float na = ad0 + adt * dt_alpha;
  na = fmaxf(na, ga.diff_alpha_min);
  const float lim = ga.diff_alpha_max * (1.f - d.avmax[i] / ga.alpha_max);
  d.adiff[i] = fminf(na, lim);
  d.ax[i] = 0.f;
  d.ay[i] = 0.f;
  d.az[i] = 0.f;
  d.udt[i] = 0.f;
  d.hdt[i] = 0.f;
  d.nf[i] = 0;
  d.mintb[i] = (int8_t)(kNumTimeBins + 1);
}

// runner_do_end_hydro_force (runner_others.c:815-870) -> hydro_end_force.
__global__ void k_end_force(DevParts d, int lo, int hi, int max_active_bin) {
  const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  if (!is_active(d.tb[i], max_active_bin)) return;
  d.hdt[i] *= d.h[i] * kDimInv;
}

__global__ void k_zero_int(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Sums the per-particle interaction counters of active particles:
// out[0] density, out[1] gradient, out[2] force.
__global__ void k_count(DevParts d, int n, int max_active_bin,
                        unsigned long long *out) {
  unsigned long long a = 0, b = 0, c = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    if (is_active(d.tb[i], max_active_bin)) {
      a += (unsigned)d.nd[i];
      b += (unsigned)d.ng[i];
      c += (unsigned)d.nf[i];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off);
    b += __shfl_down(b, off);
    c += __shfl_down(c, off);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], a);
    atomicAdd(&out[1], b);
    atomicAdd(&out[2], c);
  }
}

}  // namespace

// --------------------------------------------------------------- launchers --
namespace sphk {

static inline int nblk(int n, int b) { return (n + b - 1) / b; }

hipError_t launch_stencil(LoopKind kind, bool subset, const DevParts &d,
                          const LoopArgs &la, const int4 *tiles, int ntiles,
                          hipStream_t st) {
  if (ntiles <= 0) return hipSuccess;
  const int per = (ntiles + 7) >> 3;
  dim3 grid(per * 8), block(64);
  switch (kind) {
    case LoopKind::Density:
      if (subset)
        hipLaunchKernelGGL((k_loop_stencil<DensityLoop, true>), grid, block, 0, st, d, la, tiles, ntiles);
      else
        hipLaunchKernelGGL((k_loop_stencil<DensityLoop, false>), grid, block, 0, st, d, la, tiles, ntiles);
      break;
    case LoopKind::Gradient:
      hipLaunchKernelGGL((k_loop_stencil<GradientLoop, false>), grid, block, 0, st, d, la, tiles, ntiles);
      break;
    case LoopKind::Force:
      hipLaunchKernelGGL((k_loop_stencil<ForceLoop, false>), grid, block, 0, st, d, la, tiles, ntiles);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_task(LoopKind kind, bool subset, const DevParts &d,
                       const LoopArgs &la, const TaskArgs &ta, hipStream_t st) {
  const int nb = ta.ntiles[0] + ta.ntiles[1];
  if (nb <= 0) return hipSuccess;
  dim3 grid(nb), block(64);
  switch (kind) {
    case LoopKind::Density:
      if (subset)
        hipLaunchKernelGGL((k_loop_task<DensityLoop, true>), grid, block, 0, st, d, la, ta);
      else
        hipLaunchKernelGGL((k_loop_task<DensityLoop, false>), grid, block, 0, st, d, la, ta);
      break;
    case LoopKind::Gradient:
      hipLaunchKernelGGL((k_loop_task<GradientLoop, false>), grid, block, 0, st, d, la, ta);
      break;
    case LoopKind::Force:
      hipLaunchKernelGGL((k_loop_task<ForceLoop, false>), grid, block, 0, st, d, la, ta);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_init_density(const DevParts &d, int lo, int hi,
                               int max_active_bin, float h_max, hipStream_t st) {
  if (hi <= lo) return hipSuccess;
  hipLaunchKernelGGL(k_init_density, dim3(nblk(hi - lo, 256)), dim3(256), 0, st,
                     d, lo, hi, max_active_bin, h_max);
  return hipGetLastError();
}

hipError_t launch_ghost(const DevParts &d, const GhostArgs &ga, int lo, int hi,
                        int *redo_count, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, redo_count, 1);
  if (hi > lo)
    hipLaunchKernelGGL(k_ghost, dim3(nblk(hi - lo, 256)), dim3(256), 0, st, d,
                       ga, lo, hi, redo_count);
  return hipGetLastError();
}

hipError_t launch_extra_ghost(const DevParts &d, const GhostArgs &ga, int lo,
                              int hi, hipStream_t st) {
  if (hi <= lo) return hipSuccess;
  hipLaunchKernelGGL(k_extra_ghost, dim3(nblk(hi - lo, 256)), dim3(256), 0, st,
                     d, ga, lo, hi);
  return hipGetLastError();
}

hipError_t launch_end_force(const DevParts &d, int lo, int hi,
                            int max_active_bin, hipStream_t st) {
  if (hi <= lo) return hipSuccess;
  hipLaunchKernelGGL(k_end_force, dim3(nblk(hi - lo, 256)), dim3(256), 0, st, d,
                     lo, hi, max_active_bin);
  return hipGetLastError();
}

hipError_t launch_count(const DevParts &d, int n, int max_active_bin,
                        unsigned long long *out3, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, out3, 3);
  const int nb = n > 0 ? min(1024, nblk(n, 256)) : 1;
  hipLaunchKernelGGL(k_count, dim3(nb), dim3(256), 0, st, d, n, max_active_bin,
                     out3);
  return hipGetLastError();
}

}  // namespace sphk
