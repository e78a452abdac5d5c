// Frame statistics for the audio consumer (ml-rs/audio.rs:181-209): per-channel sums of the bytes
// and of the squared bytes of a packed 0x00RRGGBB frame, as exact integers.  The reference
// downloads the RGB24 frame and loops over it on the host; here the frame never leaves HBM
// (4 bytes per pixel read once, 48 bytes per block written).
#include <hip/hip_runtime.h>
#include "launch.h"

namespace cn {

// Each block sums a strided share of the frame (4 pixels per thread per iteration, 16-byte
// loads) and writes its six sums [s1 r, g, b, s2 r, g, b] to partial[6 * blockIdx.x ...].
__global__ void __launch_bounds__(kStatsThreads) k_render_stats(const uint4* __restrict__ frame, size_t quads,
                                                               unsigned long long* __restrict__ partial) {
  uint32_t s1[3] = {0, 0, 0};
  unsigned long long s2[3] = {0, 0, 0};
  for (size_t q = (size_t)blockIdx.x * kStatsThreads + threadIdx.x; q < quads; q += (size_t)gridDim.x * kStatsThreads) {
    const uint4 p = frame[q];
    const uint32_t px[4] = {p.x, p.y, p.z, p.w};
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        const uint32_t b = (px[j] >> (16 - 8 * ch)) & 0xFFu;  // channel 0 is the red byte
        s1[ch] += b;
        s2[ch] += b * b;
      }
  }
  __shared__ unsigned long long red[6][kStatsThreads];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    red[ch][threadIdx.x] = s1[ch];
    red[3 + ch][threadIdx.x] = s2[ch];
  }
  __syncthreads();
  for (int h = kStatsThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int j = 0; j < 6; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

hipError_t launch_render_stats(const uint32_t* frame, size_t cells, unsigned long long* partial, hipStream_t st) {
  if (cells % 4 != 0) return hipErrorInvalidValue;
  k_render_stats<<<kStatsBlocks, kStatsThreads, 0, st>>>(reinterpret_cast<const uint4*>(frame), cells / 4, partial);
  return hipGetLastError();
}

}  // namespace cn
