// Allocation-free FP64 channel mass of a caller's device state (cn_mass_fp64_noalloc).
//
// The per-block partial sums live in a file-scope device array, `scratch`, instead of an
// engine-owned buffer, so the call makes no allocation of any kind: it can run under HIP graph
// capture and in loops that must not allocate.  Each device holds its own copy of the array
// (kMaxChannels * kMassBlocks pairs of doubles, 16 KiB) until the process ends.  Its device
// address is looked up once per device with hipGetSymbolAddress (the array's name in host code
// is a host-side stand-in, not a device address) and cached per device.  Every call on a device
// shares that device's array: calls that overlap on one device race.
#include <hip/hip_runtime.h>
#include <mutex>
#include "common.h"
#include "mass.h"
#include "types.h"

namespace cn {
namespace {

constexpr int kScratchPairs = kMaxChannels * kMassBlocks;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double2 scratch[kScratchPairs];

// Partials: grid (kMassBlocks, C), block kMassThreads; channel c of the state at plane c.
__global__ void __launch_bounds__(kMassThreads)
k_mass_plane(const float* __restrict__ state, size_t plane, double2* __restrict__ part) {
  const int c = blockIdx.y;
  mass_block(state + (size_t)c * plane, plane, (int)blockIdx.x, kMassBlocks,
             part + (size_t)c * kMassBlocks + blockIdx.x);
}

// One thread per channel combines its partials in block order (the same order as the host
// combination of the engine's cn_mass_fp64).
__global__ void k_mass_final(const double2* __restrict__ part, int C, double* __restrict__ out) {
  const int c = threadIdx.x;
  if (c < C) out[c] = mass_combine(part + (size_t)c * kMassBlocks);
}

std::mutex g_mu;
double2* g_scratch[kMaxDevices] = {};

// Device address of this device's `scratch` (looked up on first use with the device current).
cn_status scratch_of(int dev, double2** out) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    cn_status st;
    if (inject_failure()) return set_error(CN_ERR_HIP, "hipGetSymbolAddress: injected failure");
    if (hip_failed(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)", &st)) return st;
    g_scratch[dev] = (double2*)p;
  }
  *out = g_scratch[dev];
  return CN_OK;
}

}  // namespace
}  // namespace cn

using namespace cn;

extern "C" cn_status cn_mass_fp64_noalloc(const float* d_state, uint32_t n, uint32_t num_channels, double* d_out,
                                          void* stream) {
  CN_TRY
  clear_error();
  if (!d_state || !d_out) return set_error(CN_ERR_INVALID_ARG, "null device pointer");
  if (num_channels == 0 || num_channels > (uint32_t)kMaxChannels)
    return set_error(CN_ERR_INVALID_ARG, "num_channels %u must be in [1, %d] (the scratch holds %d channels)",
                     num_channels, kMaxChannels, kMaxChannels);
  if (n == 0 || n > 65536) return set_error(CN_ERR_INVALID_ARG, "grid side %u must be in [1, 65536]", n);
  int dev = -1;
  CN_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(CN_ERR_UNSUPPORTED, "device ordinal %d", dev);
  double2* part = nullptr;
  const cn_status st = scratch_of(dev, &part);
  if (st != CN_OK) return st;
  const hipStream_t s = (hipStream_t)stream;
  // Every partial read by k_mass_final is written by k_mass_plane first: nothing to clear.
  k_mass_plane<<<dim3(kMassBlocks, num_channels), kMassThreads, 0, s>>>(d_state, (size_t)n * n, part);
  k_mass_final<<<1, kMaxChannels, 0, s>>>(part, (int)num_channels, d_out);
  CN_HIP(hipGetLastError());
  return CN_OK;
  CN_CATCH
}
