// sph_scratch.hip -- the fixed device scratch of the allocation-free batched
// pass (swsph_pass_batched_noalloc).
//
// The per-pass scratch of the batched pass -- the work queue (particle
// flags, the scans, the compact particle list, the tile table, its count
// and the per-XCD queue heads) and the ghost / interaction counters -- is
// carved from one file-scope device array instead of engine allocations.
// Every process that loads the library holds the array on each GPU it uses
// until it exits.  The array's name is only a host-side handle: host code
// uses the device address that hipGetSymbolAddress returns for the current
// device, resolved once per device (at engine creation) and kept here.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sph_kernels.h"

namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the scratch of a pass over n particles on nc cells, in the order
// noalloc_scratch carves them.
constexpr size_t scratch_bytes(size_t n, size_t nc) {
  const size_t tcap = n / kMinTile + nc + 1;
  return align256(4 * n)                   // flag
         + align256(4 * (n + 1))           // scan
         + align256(4 * (n / 1024 + 2))    // blk
         + align256(4 * nc)                // cnt
         + align256(4 * (nc + 1))          // cscan
         + align256(4 * (nc / 1024 + 2))   // cblk
         + align256(4 * n)                 // alist
         + 2 * align256(4 * tcap)          // tile_start, tile_cell
         + align256(4)                     // ntiles
         + align256(4 * 8 * 64)            // heads
         + align256(8 * kGhostCounters)    // ghost counters
         + align256(8 * 3);                // interaction counters
}

constexpr size_t kScratchBytes =
    scratch_bytes((size_t)kNoAllocMaxParts, (size_t)kNoAllocMaxCells);

constexpr int kMaxDevices = 64;
std::mutex g_mtx;
void *g_ptr[kMaxDevices] = {};

}  // namespace

// The scratch itself (about 30 MB for the limits in sph_kernels.h).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace sphk {

size_t noalloc_scratch_bytes() { return kScratchBytes; }

hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mtx);
  if (g_ptr[device]) return hipSuccess;
  int cur = -1;
  hipError_t err = hipGetDevice(&cur);
  if (err != hipSuccess) return err;
  if (cur != device) return hipErrorInvalidDevice;  // caller makes it current
  void *p = nullptr;
  err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (err != hipSuccess) return err;
  if (!p) return hipErrorInvalidSymbol;
  g_ptr[device] = p;
  return hipSuccess;
}

hipError_t noalloc_scratch(int device, int64_t n, int ncells, TileScratch *ts,
                           size_t *tile_cap, unsigned long long **redo,
                           unsigned long long **cnt) {
  if (n < 0 || n > kNoAllocMaxParts || ncells < 0 ||
      ncells > kNoAllocMaxCells)
    return hipErrorInvalidValue;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  unsigned char *base;
  {
    std::lock_guard<std::mutex> lk(g_mtx);
    base = (unsigned char *)g_ptr[device];
  }
  if (!base) return hipErrorNotInitialized;
  const size_t nn = (size_t)n, nc = (size_t)ncells;
  const size_t tcap = nn / kMinTile + nc + 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char *p = base + off;
    off += align256(bytes);
    return p;
  };
  ts->flag = (int *)take(4 * nn);
  ts->scan = (int *)take(4 * (nn + 1));
  ts->blk = (int *)take(4 * (nn / 1024 + 2));
  ts->cnt = (int *)take(4 * nc);
  ts->cscan = (int *)take(4 * (nc + 1));
  ts->cblk = (int *)take(4 * (nc / 1024 + 2));
  ts->alist = (int *)take(4 * nn);
  ts->tile_start = (int *)take(4 * tcap);
  ts->tile_cell = (int *)take(4 * tcap);
  ts->ntiles = (int *)take(4);
  ts->heads = (int *)take(4 * 8 * 64);
  *redo = (unsigned long long *)take(8 * kGhostCounters);
  *cnt = (unsigned long long *)take(8 * 3);
  if (off != scratch_bytes(nn, nc) || off > kScratchBytes)
    return hipErrorInvalidValue;
  *tile_cap = tcap;
  return hipSuccess;
}

}  // namespace sphk
