// gz_noalloc.hip -- Sketch::update over a batch (src/sketch.cpp:69-83) for callers that cannot
// allocate device memory: code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc.  It runs the same update kernel as gz_sketch_update_batch; the kernel's work queue of
// tile descriptors lives in a file-scope device array instead of a buffer the library allocates,
// and the indices are read where the caller keeps them in device memory (nothing is staged).
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gz_hip.h"
#include "gz_runtime.hpp"

namespace gz {
namespace {

// One call's tile descriptors (one wave of the update kernel per 1024 indices).  Each GPU holds its
// own copy of this array until the process ends.  Every call on a GPU shares that GPU's copy and
// overwrites the descriptors it uses before the kernel reads them (nothing must start at zero), so
// calls that overlap on one GPU -- from several threads, or on several sketches -- race.
constexpr uint32_t kNoAllocTile = 1024;
constexpr size_t kNoAllocTiles = 65536;  // 1 MiB of descriptors
static __device__ __align__(256) Tile scratch[kNoAllocTiles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mtx;
Tile* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the first call there.  On
// the host, `scratch` names a host-side stand-in, never the device memory, so only this address is
// passed to kernels.
Tile* device_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw GzError(GZ_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mtx);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw GzError(GZ_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<Tile*>(p);
  }
  return g_scratch[dev];
}

// Makes `dev` current for the call and restores the caller's device afterwards.
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    GZ_HIP(hipGetDevice(&prev));
    if (prev != dev) GZ_HIP(hipSetDevice(dev));
    else prev = -1;
  }
  ~DeviceScope() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// Tile t covers indices [t * tile, min(count, (t + 1) * tile)) of the one sketch (row 0); with more
// than one tile the write-back merges with atomics, as gz_sketch_update_batch's tiles do.
__global__ void k_seq_tiles(Tile* __restrict__ tiles, uint32_t nt, uint32_t count, uint32_t tile) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    const uint32_t b = t * tile;
    tiles[t] = Tile{0u, b, count - b < tile ? count - b : tile, nt > 1 ? 1u : 0u};
  }
}

}  // namespace
}  // namespace gz

extern "C" int gz_sketch_update_batch_device_noalloc(gz_sketch* s, const uint64_t* idx_device, size_t count) {
  try {
    if (!s || (!idx_device && count)) throw gz::GzError(GZ_ERR_INVALID_ARG, "null argument");
    if (count > gz::kNoAllocTiles * gz::kNoAllocTile)
      throw gz::GzError(GZ_ERR_INVALID_ARG, "count above the no-allocation limit of 2^26 indices per call");
    if (count) {
      gz::GzSketch& S = *s->s;
      gz::DeviceScope scope(S.dev);
      gz::Tile* tiles = gz::device_scratch(S.dev);
      const uint32_t nt = (uint32_t)((count + gz::kNoAllocTile - 1) / gz::kNoAllocTile);
      hipLaunchKernelGGL(gz::k_seq_tiles, dim3((nt + 255) / 256), dim3(256), 0, S.st, tiles, nt, (uint32_t)count,
                         gz::kNoAllocTile);
      gz::launch_update_tiles(S.gb, tiles, nullptr, nt, 0, nullptr, idx_device, S.slab.p, S.geo, S.col_bf.p, S.st);
      GZ_HIP(hipGetLastError());
      GZ_HIP(hipStreamSynchronize(S.st));
    }
    gz::clear_last_error();
    return GZ_OK;
  } catch (const std::exception& e) {
    return gz::status_of(e);
  }
}
