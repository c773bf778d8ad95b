// sph_step.hip -- what changes between two neighbour passes, on the device:
// binning of the particles into the top-level grid (the particle sort of
// space_rebuild plus the in-cell order of the chunk engine) and the time
// integration of the gas (drift, kicks, time steps).
//
// Binning: every particle gets the 64-bit key (cell << 30) | morton, where
// morton interleaves 10 bits per axis of the cell-relative position (x in
// the highest bit of each triple).  A stable LSD radix sort (rocPRIM) of the
// keys with the current slot as value gives the new order; cells start
// where their key range starts.  With the caller's index as the value this
// is exactly the host rule "sort by (cell, morton, caller index)".  With
// the time-bin order (kBinBits > 0) the key is (cell << 36) | (time bin <<
// 30) | morton: inside a cell the particles of the shortest steps come
// first, so the particles active at any step form a prefix of each cell
// (their fields are contiguous runs instead of one line per particle).
//
// Time integration restates, for gas without gravity and without
// cosmology:
//   drift_part (src/drift.h:138) + hydro_predict_extra
//     (src/hydro/SPHENIX/hydro.h:1029-1078),
//   kick_part (src/kick.h) + hydro_kick_extra (hydro.h:1120-1151), with the
//     half-kick intervals of runner_do_kick1 / runner_do_kick2
//     (src/runner_time_integration.c:87, 360) and
//     hydro_reset_predicted_values (hydro.h:983-1006) after kick2,
//   get_part_timestep (src/timestep.h) with hydro_compute_timestep
//     (hydro.h:475-487) and make_integer_timestep (src/timestep.h:49-89),
//     on the integer time line of src/timeline.h.
// Floating-point expressions keep the reference's operand order and types
// (float fields, double time intervals); contraction is off so that each
// operation rounds as written.
#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>

#include "sph_device.h"
#include "sph_kernels.h"
#include "sph_loops.h"

using namespace sph;
using namespace sphl;

namespace {

constexpr int kBlock = 256;
// Grid cap of the kernels that reduce into one counter: a grid-stride loop
// over the particles, one atomic per block.
constexpr int kReduceBlocks = 4096;
constexpr long long kMaxNrTimesteps = 1LL << (kNumTimeBins + 1);
constexpr int kNeighbourMaxDeltaBin = 2;  // time_bin_neighbour_max_delta_bin

int blocks(int n) { return (n + kBlock - 1) / kBlock; }
int reduce_blocks(int n) { return std::max(1, std::min(blocks(n), kReduceBlocks)); }

// ---------------------------------------------------------------- binning --
__device__ __forceinline__ unsigned long long spread10(unsigned int v) {
  unsigned long long x = v & 0x3ffu;
  x = (x | (x << 16)) & 0x030000FFull;
  x = (x | (x << 8)) & 0x0300F00Full;
  x = (x | (x << 4)) & 0x030C30C3ull;
  x = (x | (x << 2)) & 0x09249249ull;
  return x;
}

// Key of particle i from its position (component k at p[k][stride * i]),
// the cell rule of space_parts_get_cell_index (space_cell_index.c:114-128):
// (int)(x * iwidth) with iwidth = 1 / (dim / cdim), clamped to the grid;
// 10 Morton bits per axis from (x * iwidth - cell) * 1024.
__global__ void k_bin_keys(const double *px, const double *py,
                           const double *pz, int stride, GridDesc g,
                           double iw0, double iw1, double iw2, int n,
                           unsigned long long *key, int *val,
                           const int8_t *tb, int bin_bits) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x[3] = {px[(size_t)stride * i], py[(size_t)stride * i],
                       pz[(size_t)stride * i]};
  const double iw[3] = {iw0, iw1, iw2};
  int c[3];
  unsigned long long m = 0ull;
  for (int k = 0; k < 3; k++) {
    const double s = x[k] * iw[k];
    int v = (int)s;
    v = v < 0 ? 0 : (v >= g.nc[k] ? g.nc[k] - 1 : v);
    c[k] = v;
    const double f = (s - (double)v) * 1024.0;
    int q = (int)f;
    q = q < 0 ? 0 : (q > 1023 ? 1023 : q);
    m |= spread10((unsigned int)q) << (2 - k);
  }
  const unsigned long long cid =
      (unsigned long long)((c[0] * g.nc[1] + c[1]) * g.nc[2] + c[2]);
  unsigned long long b = 0ull;
  if (bin_bits) {  // inhibited particles (and any bin past the field) last
    const int t = tb[i];
    b = (unsigned long long)(t < 0 ? 0 : (t > 63 ? 63 : t));
  }
  key[i] = (cid << (30 + bin_bits)) | (b << 30) | m;
  val[i] = i;
}

// Periodic wrap of the positions into [0, dim) (box_wrap of space_rebuild;
// a particle moves by less than a box length between two binnings).
__global__ void k_wrap(DevParts d, GridDesc g, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double *p[3] = {d.px, d.py, d.pz};
  for (int k = 0; k < 3; k++) {
    double x = p[k][i];
    if (x >= g.dim[k])
      x -= g.dim[k];
    else if (x < 0.0)
      x += g.dim[k];
    // a position wrapped onto the upper face by rounding goes to 0
    // (space_cell_index.c:113-122)
    if (x == g.dim[k]) x = 0.0;
    p[k][i] = x;
  }
}

// Cell of every sorted slot, and the first slot of every cell (cells
// without particles start where the next non-empty one does).
__global__ void k_cell_of(const unsigned long long *key, int n, int *cell,
                          int shift) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) cell[k] = (int)(key[k] >> shift);
}

__global__ void k_cell_start(const unsigned long long *key, int n, int ncells,
                             int *cs, int shift) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncells) return;
  const unsigned long long t = (unsigned long long)c << shift;
  int lo = 0, hi = n;  // first slot with key >= t
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (key[mid] < t)
      lo = mid + 1;
    else
      hi = mid;
  }
  cs[c] = lo;
}

template <class T>
__global__ void k_gather(T *dst, const T *src, int stride, int comp,
                         const int *idx, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[(size_t)stride * idx[k] + comp];
}

// dst[stride * idx[k] + comp] = src[k]: back to the caller's order.
template <class T>
__global__ void k_scatter(T *dst, const T *src, int stride, int comp,
                          const int *idx, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[(size_t)stride * idx[k] + comp] = src[k];
}

template <class T>
__global__ void k_fill(T *dst, T v, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = v;
}

// ---------------------------------------------------------- time line ----
__device__ __forceinline__ long long integer_timestep(int bin) {
  return bin <= 0 ? 0LL : (1LL << (bin + 1));
}

__device__ __forceinline__ int time_bin_of(long long dti) {
  return (8 * (int)sizeof(long long) - 2) - __clzll(dti);
}

__device__ __forceinline__ long long time_begin(long long ti, int bin) {
  const long long dti = integer_timestep(bin);
  return dti == 0 ? 0LL : dti * ((ti - 1) / dti);
}

__device__ __forceinline__ long long time_end(long long ti, int bin) {
  const long long dti = integer_timestep(bin);
  if (dti == 0) return 0LL;
  const long long mod = ti % dti;
  return mod == 0 ? ti : ti - mod + dti;
}

// 4th-order expansion of exp(x) (src/approx_math.h:35-37).
__device__ __forceinline__ float approx_expf(float x) {
#pragma clang fp contract(off)
  return 1.f + x * (1.f + x * (0.5f + x * (1.f / 6.f + 1.f / 24.f * x)));
}

__device__ __forceinline__ float gas_pressure(float rho, float u) {
#pragma clang fp contract(off)
  return kGasGm1 * u * rho;
}

__device__ __forceinline__ float gas_soundspeed(float rho, float P) {
#pragma clang fp contract(off)
  return sqrtf(kGasGamma * P / rho);
}

__device__ __forceinline__ float wave_max_u(float v) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}

__device__ __forceinline__ long long wave_min_ll(long long v) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const long long o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}

// drift_part + hydro_predict_extra for the particles [lo, hi), and the
// largest distance by which a drifted particle lies outside its cell's box
// (max-reduced into *dmax as float bits; the value is non-negative).
__device__ __forceinline__ float drift_one(const DevParts &d, const DriftArgs &a,
                                           int i) {
#pragma clang fp contract(off)
  const double dt = a.dt_drift;
  const double x = d.px[i] + (double)d.vfx[i] * dt;
  const double y = d.py[i] + (double)d.vfy[i] * dt;
  const double z = d.pz[i] + (double)d.vfz[i] * dt;
  d.px[i] = x;
  d.py[i] = y;
  d.pz[i] = z;
  // predicted velocities (hydro terms)
  d.vx[i] = (float)((double)d.vx[i] + (double)d.ax[i] * a.dt_kick_hydro);
  d.vy[i] = (float)((double)d.vy[i] + (double)d.ay[i] * a.dt_kick_hydro);
  d.vz[i] = (float)((double)d.vz[i] + (double)d.az[i] * a.dt_kick_hydro);
  // hydro_predict_extra takes its intervals as float
  const float dt_drift = (float)a.dt_drift;
  const float dt_therm = (float)a.dt_therm;
  float u = d.u[i] + d.udt[i] * dt_therm;
  float h = d.h[i];
  const float h_inv = 1.f / h;
  const float w1 = d.hdt[i] * h_inv * dt_drift;
  h *= fabsf(w1) < 0.2f ? approx_expf(w1) : expf(w1);
  const float w2 = -kDim * w1;
  float rho = d.rho[i];
  rho *= fabsf(w2) < 0.2f ? approx_expf(w2) : expf(w2);
  // entropy floor: none (floor u = 0); absolute minimum
  u = fmaxf(u, 0.f);
  u = fmaxf(u, a.min_u);
  const float P = gas_pressure(rho, u);
  const float c = gas_soundspeed(rho, P);
  d.u[i] = u;
  d.h[i] = h;
  d.rho[i] = rho;
  d.P[i] = P;
  d.c[i] = c;
  d.vsig[i] = fmaxf(d.vsig[i], 2.f * c);
  // distance outside the box of the particle's cell
  const int cid = d.cell[i];
  const int cz = cid % a.g.nc[2], cy = (cid / a.g.nc[2]) % a.g.nc[1];
  const int cx = cid / (a.g.nc[2] * a.g.nc[1]);
  const int cc[3] = {cx, cy, cz};
  const double p[3] = {x, y, z};
  double o = 0.0;
  for (int k = 0; k < 3; k++) {
    const double w = a.g.dim[k] / a.g.nc[k];
    const double lo = (double)cc[k] * w;
    o = fmax(o, fmax(lo - p[k], p[k] - (lo + w)));
  }
  // rounded up so that the bound stays a bound
  float out = (float)o;
  if ((double)out < o) out = nextafterf(out, 1e30f);
  return out;
}

__global__ __launch_bounds__(kBlock) void k_drift(DevParts d, DriftArgs a,
                                                  unsigned int *dmax) {
  __shared__ float red[kBlock / 64];
  const int i = a.lo + blockIdx.x * blockDim.x + threadIdx.x;
  float out = 0.f;
  if (i < a.hi && d.tb[i] != kBinInhibited) out = drift_one(d, a, i);
  out = wave_max_u(out);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = out;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = red[0];
    for (int k = 1; k < kBlock / 64; k++) m = fmaxf(m, red[k]);
    // a plain read first: only a block that raises the maximum pays an
    // atomic (one per block on one address serialised large drifts)
    if (m > 0.f &&
        __float_as_uint(m) >
            __hip_atomic_load(dmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(dmax, __float_as_uint(m));
  }
}

// kick_part + hydro_kick_extra over [lo, hi): mode 1 the first half-kick of
// the particles starting a step at ti_current (runner_do_kick1), mode 2 the
// second half-kick of the active particles followed by
// hydro_reset_predicted_values (runner_do_kick2).
__device__ __forceinline__ void kick_one(const DevParts &d, const KickArgs &a,
                                         int i, int tb) {
#pragma clang fp contract(off)
  if (tb == kBinInhibited) return;
  const long long dti = integer_timestep(tb);
  long long ti_begin, ti_end;
  if (a.mode == 1) {
    ti_begin = time_begin(a.ti_current + 1, tb);
    if (ti_begin != a.ti_current) return;  // not starting a step
    ti_end = ti_begin + dti / 2;
  } else {
    if (tb > a.max_active_bin) return;  // part_is_active
    ti_begin = time_begin(a.ti_current, tb) + dti / 2;
    ti_end = ti_begin + dti / 2;
  }
  const double dt = (double)(ti_end - ti_begin) * a.time_base;
  float vfx = (float)((double)d.vfx[i] + (double)d.ax[i] * dt);
  float vfy = (float)((double)d.vfy[i] + (double)d.ay[i] * dt);
  float vfz = (float)((double)d.vfz[i] + (double)d.az[i] * dt);
  d.vfx[i] = vfx;
  d.vfy[i] = vfy;
  d.vfz[i] = vfz;
  // hydro_kick_extra (dt_therm as float)
  const float delta_u = d.udt[i] * (float)dt;
  float ufull = d.ufull[i];
  ufull = fmaxf(ufull + delta_u, 0.5f * ufull);
  const float energy_min = fmaxf(a.min_u, 0.f);
  if (ufull < energy_min) {
    ufull = energy_min;
    d.udt[i] = 0.f;
  }
  d.ufull[i] = ufull;
  if (a.mode == 2) {
    d.vx[i] = vfx;
    d.vy[i] = vfy;
    d.vz[i] = vfz;
    d.u[i] = ufull;
    const float rho = d.rho[i];
    const float P = gas_pressure(rho, ufull);
    const float c = gas_soundspeed(rho, P);
    d.P[i] = P;
    d.c[i] = c;
    d.vsig[i] = fmaxf(d.vsig[i], 2.f * c);
  }
}

__global__ void k_kick(DevParts d, KickArgs a) {
  for_runs(a.lo, a.hi, d.tb, [&](int i, int8_t tb) { kick_one(d, a, i, tb); });
}

// make_integer_timestep (src/timestep.h:49-89); 0 flags a step of size 0.
__device__ long long make_integer_timestep(float new_dt, int old_bin,
                                           int min_ngb_bin, long long ti_current,
                                           double time_base_inv) {
#pragma clang fp contract(off)
  long long new_dti = (long long)((double)new_dt * time_base_inv);
  if (new_dti <= 0) return 0;
  int new_bin = time_bin_of(new_dti);
  new_bin = min(new_bin, min_ngb_bin + kNeighbourMaxDeltaBin);
  new_dti = integer_timestep(new_bin);
  const long long current_dti = integer_timestep(old_bin);
  const long long ti_end = time_end(ti_current, old_bin);
  if (old_bin > 0) new_dti = min(new_dti, 2 * current_dti);
  long long dti_timeline = kMaxNrTimesteps;
  while (new_dti < dti_timeline) dti_timeline /= 2;
  new_dti = dti_timeline;
  if (new_dti > current_dti) {
    if ((kMaxNrTimesteps - ti_end) % new_dti > 0) new_dti = current_dti;
  }
  return new_dti;
}

// runner_do_timestep over [lo, hi): new time bins of the active particles
// (kept as they are with `fixed`), and the next synchronisation point: the
// smallest end of step over all particles, min-reduced into *ti_end_min.
// err[0] counts particles whose step falls below dt_min or to zero, err[1]
// the active particles.
__device__ __forceinline__ long long timestep_one(const DevParts &d,
                                                  const StepArgs &a, int i,
                                                  unsigned int *err,
                                                  bool &active) {
#pragma clang fp contract(off)
  long long ti_end = kMaxNrTimesteps;
  active = false;
  if (d.tb[i] != kBinInhibited) {
    const int tb = d.tb[i];
    active = tb <= a.max_active_bin;
    if (active) {
      long long new_dti;
      if (a.fixed) {
        new_dti = integer_timestep(tb);
        if (new_dti == 0) {  // a step of size 0 never ends
          atomicAdd(err, 1u);
          new_dti = 1;
        }
      } else {
        const float h = d.h[i];
        // hydro_compute_timestep; cosmo->a and a_factor_sound_speed are
        // doubles (1 without cosmology), so the quotient is formed in double
        const float dt_cfl = (float)((double)(2.f * kGamma * a.cfl) * 1.0 *
                                     (double)h / (1.0 * (double)d.vsig[i]));
        float new_dt = dt_cfl;
        const float hdt = d.hdt[i];
        const float dt_h_change =
            hdt != 0.f ? fabsf(a.log_max_h_change * h / hdt) : 3.402823466e38f;
        new_dt = fminf(new_dt, dt_h_change);
        new_dt = (float)fmin((double)new_dt, a.dt_max);
        if ((double)new_dt < a.dt_min) {
          atomicAdd(err, 1u);
          new_dti = integer_timestep(tb);
        } else {
          new_dti = make_integer_timestep(new_dt, tb, d.mintb[i], a.ti_current,
                                          a.time_base_inv);
          if (new_dti == 0) {
            atomicAdd(err, 1u);
            new_dti = integer_timestep(tb);
          } else {
            d.tb[i] = (int8_t)time_bin_of(new_dti);
          }
        }
      }
      ti_end = a.ti_current + new_dti;
    } else {
      ti_end = time_end(a.ti_current, tb);
    }
  }
  return ti_end;
}

__global__ __launch_bounds__(kBlock) void k_timestep(DevParts d, StepArgs a,
                                                     unsigned long long *ti_end_min,
                                                     unsigned int *err) {
  __shared__ long long red[kBlock / 64];
  long long ti_end = kMaxNrTimesteps;
  int nloc = 0;
  for (int i = a.lo + blockIdx.x * blockDim.x + threadIdx.x; i < a.hi;
       i += gridDim.x * blockDim.x) {
    bool act;
    const long long t = timestep_one(d, a, i, err, act);
    ti_end = t < ti_end ? t : ti_end;
    nloc += act ? 1 : 0;
  }
  ti_end = wave_min_ll(ti_end);
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) nloc += __shfl_xor(nloc, off);
  __shared__ int rn[kBlock / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = ti_end;
    rn[w] = nloc;
  }
  __syncthreads();
  int nact = 0;
  for (int k = 0; k < kBlock / 64; k++) nact += rn[k];
  if (threadIdx.x == 0) {
    long long m = red[0];
    for (int k = 1; k < kBlock / 64; k++) m = red[k] < m ? red[k] : m;
    atomicMin(ti_end_min, (unsigned long long)m);
    if (nact) atomicAdd(err + 1, (unsigned int)nact);
  }
}

}  // namespace

namespace sphk {

hipError_t launch_bin_keys(const double *px, const double *py, const double *pz,
                           int stride, const GridDesc &g, const double inv_w[3],
                           int n, unsigned long long *key, int *val,
                           hipStream_t st, const int8_t *tb) {
  if (n <= 0) return hipSuccess;
  k_bin_keys<<<blocks(n), kBlock, 0, st>>>(px, py, pz, stride, g, inv_w[0],
                                           inv_w[1], inv_w[2], n, key, val, tb,
                                           tb ? kBinBits : 0);
  return hipGetLastError();
}

hipError_t sort_pairs(void *temp, size_t &bytes, const unsigned long long *kin,
                      unsigned long long *kout, const int *vin, int *vout,
                      int n, int bits, hipStream_t st) {
  return rocprim::radix_sort_pairs(temp, bytes, kin, kout, vin, vout,
                                   (unsigned int)n, 0u, (unsigned int)bits, st);
}

hipError_t launch_wrap(const DevParts &d, const GridDesc &g, int n,
                       hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_wrap<<<blocks(n), kBlock, 0, st>>>(d, g, n);
  return hipGetLastError();
}

hipError_t launch_cell_bounds(const unsigned long long *key, int n, int ncells,
                              int *cell_start, int *cell, hipStream_t st,
                              int bin_bits) {
  const int shift = 30 + bin_bits;
  if (n > 0) k_cell_of<<<blocks(n), kBlock, 0, st>>>(key, n, cell, shift);
  k_cell_start<<<blocks(ncells + 1), kBlock, 0, st>>>(key, n, ncells,
                                                      cell_start, shift);
  return hipGetLastError();
}

template <class T>
hipError_t launch_gather(T *dst, const T *src, int stride, int comp,
                         const int *idx, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_gather<T><<<blocks(n), kBlock, 0, st>>>(dst, src, stride, comp, idx, n);
  return hipGetLastError();
}
template hipError_t launch_gather<double>(double *, const double *, int, int,
                                          const int *, int, hipStream_t);
template hipError_t launch_gather<float>(float *, const float *, int, int,
                                         const int *, int, hipStream_t);
template hipError_t launch_gather<int>(int *, const int *, int, int,
                                       const int *, int, hipStream_t);
template hipError_t launch_gather<int8_t>(int8_t *, const int8_t *, int, int,
                                          const int *, int, hipStream_t);

template <class T>
hipError_t launch_scatter(T *dst, const T *src, int stride, int comp,
                          const int *idx, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_scatter<T><<<blocks(n), kBlock, 0, st>>>(dst, src, stride, comp, idx, n);
  return hipGetLastError();
}
template hipError_t launch_scatter<double>(double *, const double *, int, int,
                                           const int *, int, hipStream_t);
template hipError_t launch_scatter<float>(float *, const float *, int, int,
                                          const int *, int, hipStream_t);
template hipError_t launch_scatter<int>(int *, const int *, int, int,
                                        const int *, int, hipStream_t);
template hipError_t launch_scatter<int8_t>(int8_t *, const int8_t *, int, int,
                                           const int *, int, hipStream_t);

hipError_t launch_fill_i8(int8_t *p, int8_t v, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  k_fill<int8_t><<<blocks(n), kBlock, 0, st>>>(p, v, n);
  return hipGetLastError();
}

hipError_t launch_drift(const DevParts &d, const DriftArgs &a,
                        unsigned int *dmax, hipStream_t st) {
  if (a.hi <= a.lo) return hipSuccess;
  k_drift<<<blocks(a.hi - a.lo), kBlock, 0, st>>>(d, a, dmax);
  return hipGetLastError();
}

hipError_t launch_kick(const DevParts &d, const KickArgs &a, hipStream_t st) {
  if (a.hi <= a.lo) return hipSuccess;
  k_kick<<<run_blocks(a.hi - a.lo, kBlock), kBlock, 0, st>>>(d, a);
  return hipGetLastError();
}

hipError_t launch_timestep(const DevParts &d, const StepArgs &a,
                           unsigned long long *ti_end_min, unsigned int *err,
                           hipStream_t st) {
  if (a.hi <= a.lo) return hipSuccess;
  k_timestep<<<reduce_blocks(a.hi - a.lo), kBlock, 0, st>>>(d, a, ti_end_min,
                                                          err);
  return hipGetLastError();
}

}  // namespace sphk
