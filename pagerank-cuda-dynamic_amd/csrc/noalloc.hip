// noalloc.hip -- prh_partition_noalloc: the degree partition of prh_partition with its scratch in a fixed device
// array instead of context-allocated buffers, for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The partition's scratch is the per-tile class counts and offsets (3 uint32 per tile of kPartitionTile vertices)
// and the three class sizes. They live in `scratch`, a file-scope device array of kNoAllocScratchBytes, so the
// entry point runs for graphs of up to kNoAllocMaxVertices vertices. Each GPU holds its own copy of the array (at
// its own address) until the process ends. Host code never uses the array's name as an address: the context's
// set-up resolves its device address once per device with hipGetSymbolAddress (noalloc_setup), and that pointer is
// what the kernels receive.
//
// Every call on one GPU shares that GPU's array and leaves its contents behind; each call rewrites all of the
// scratch it reads (the count pass writes every tile's counts before the scan reads them), so nothing needs
// clearing. Calls that overlap on one GPU (two streams, two contexts or two threads at once) race on the array and
// must be ordered by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "capi_internal.hpp"

namespace {

constexpr uint32_t kNoAllocMaxVertices = 1u << 28;
constexpr size_t kCountWords = 64;  // the 3 class sizes, padded to 256 bytes so the tiles stay aligned
constexpr size_t kTileWords = 3 * size_t((kNoAllocMaxVertices + prh::kPartitionTile - 1) / prh::kPartitionTile);
constexpr size_t kNoAllocScratchWords = kCountWords + kTileWords;
constexpr size_t kNoAllocScratchBytes = kNoAllocScratchWords * 4;  // 1.5 MiB + 256 bytes
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

namespace {

std::mutex g_lock;
uint32_t *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device, once resolved

// out[0] = n ? counts[0] : 0 (the partition point, into the caller's device word).
__global__ void store_point_kernel(const uint32_t *counts, uint32_t n, uint32_t *out) {
  if (threadIdx.x == 0) out[0] = n ? counts[0] : 0u;
}

}  // namespace

namespace prhc {

// Resolves the device address of `scratch` on the context's device (the current device) once per device and keeps
// it in the context. Called by prh_context_create after hipSetDevice.
int noalloc_setup(prh_context *c) {
  if (c->device < 0 || c->device >= kMaxDevices)
    return fail(c, PRH_ERROR_UNSUPPORTED, "device %d: the no-allocation scratch table holds %d devices", c->device,
                kMaxDevices);
  std::lock_guard<std::mutex> guard(g_lock);
  if (!g_scratch[c->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(c, PRH_ERROR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    g_scratch[c->device] = static_cast<uint32_t *>(p);
  }
  c->noalloc_scratch = g_scratch[c->device];
  return PRH_SUCCESS;
}

}  // namespace prhc

using namespace prhc;

extern "C" {

int prh_partition_noalloc(prh_context *c, uint32_t n, const uint32_t *offsets, uint32_t switch_degree,
                          uint32_t *order, uint32_t *partition_point) {
  if (!c) return PRH_ERROR_INVALID_ARGUMENT;
  if (!partition_point || (n && (!offsets || !order))) return fail(c, PRH_ERROR_INVALID_ARGUMENT, "null argument");
  if (n > kNoAllocMaxVertices)
    return fail(c, PRH_ERROR_INVALID_ARGUMENT,
                "prh_partition_noalloc: %u vertices exceed its fixed scratch (at most %u vertices, %zu bytes)", n,
                kNoAllocMaxVertices, kNoAllocScratchBytes);
  if (!c->noalloc_scratch) return fail(c, PRH_ERROR_DEVICE, "no-allocation scratch not resolved for this device");
  PRH_HIP(c, hipSetDevice(c->device));
  uint32_t *counts = c->noalloc_scratch;
  uint32_t *tiles = c->noalloc_scratch + kCountWords;
  if (n) PRH_HIP(c, prh::launch_partition(offsets, n, switch_degree, 0xFFFFFFFFu, order, tiles, counts, c->stream));
  store_point_kernel<<<1, 64, 0, c->stream>>>(counts, n, partition_point);
  PRH_HIP(c, hipGetLastError());
  return PRH_SUCCESS;
}

}  // extern "C"
