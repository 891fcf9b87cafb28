// faces_noalloc.hip -- the staged step for callers that cannot allocate.
//
// faces_share_noalloc runs the staged step of one rank (pack straight into
// the receive slot of the opposite direction, interior sums, surface sums:
// faces_share on FACES_PATH_STAGED_DIRECT) without a plan: its receive
// messages live in `scratch`, a device array of FACES_NOALLOC_SCRATCH_BYTES
// that the library holds on every GPU it runs on until the process ends.
//
// In host code the name `scratch` denotes a host-side stand-in, not device
// memory.  The array's device address is taken once per device, with that
// device current, through hipGetSymbolAddress, and that address -- never the
// array's name -- is what the kernels receive.  Every call on one GPU shares
// that GPU's array; calls that overlap on one GPU (other streams, other
// threads) race on it.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

#include "faces_common.hpp"

static constexpr size_t kScratchBytes = FACES_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size");

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of `scratch` on the current device, looked up on the
// first call on that device.
int scratch_on_current_device(double **out)
{
  int dev = 0;
  FACES_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return faces::fail(FACES_ERR_NODEV, "device %d outside [0,%d)", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return faces::hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return faces::fail(FACES_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
    g_scratch[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch[dev];
  return FACES_OK;
}

}  // namespace

extern "C" int faces_share_noalloc(double *u, int rank, int lx, int ly, int lz, int mx, int my, int mz,
                                   int n, int boundary, long plane_stride, int compute, void *stream)
{
  if (!u) return faces::fail(FACES_ERR_ARG, "null field pointer");
  double *arena = nullptr;
  const int st = scratch_on_current_device(&arena);
  if (st) return st;
  return faces::share_staged_on_arena(u, rank, lx, ly, lz, mx, my, mz, n, boundary, plane_stride, compute, arena,
                                      kScratchBytes, "faces_share_noalloc", (hipStream_t)stream);
}
