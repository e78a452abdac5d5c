// FP64 Neumaier-compensated channel mass: per-block partial sums on the device and the final
// combination over the blocks, shared by every entry point that reports mass so that they all
// give bitwise the same sums.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstddef>

namespace cn {

constexpr int kMassThreads = 256;
constexpr int kMassBlocks = 64;  // partial sums per channel

__host__ __device__ __forceinline__ void neumaier_add(double& s, double& comp, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x)) comp += (s - t) + x;
  else comp += (x - t) + s;
  s = t;
}

#ifdef __HIPCC__
// Block `blk` of `nblk` (kMassThreads threads) sums plane[i] for i = blk*kMassThreads + tid +
// j*nblk*kMassThreads, then reduces the block's threads in a fixed tree; thread 0 writes the
// (sum, compensation) pair to *dst.  Every block of a grid must call it (it has barriers).
__device__ __forceinline__ void mass_block(const float* __restrict__ plane, size_t n, int blk, int nblk,
                                           double2* dst) {
  double sum = 0.0, comp = 0.0;
  for (size_t i = (size_t)blk * kMassThreads + threadIdx.x; i < n; i += (size_t)nblk * kMassThreads)
    neumaier_add(sum, comp, (double)plane[i]);
  __shared__ double ss[kMassThreads], sc[kMassThreads];
  ss[threadIdx.x] = sum;
  sc[threadIdx.x] = comp;
  __syncthreads();
  for (int h = kMassThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      double a = ss[threadIdx.x], ca = sc[threadIdx.x] + sc[threadIdx.x + h];
      neumaier_add(a, ca, ss[threadIdx.x + h]);
      ss[threadIdx.x] = a;
      sc[threadIdx.x] = ca;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *dst = make_double2(ss[0], sc[0]);
}
#endif

// Mass of one channel from its kMassBlocks partial pairs (sum, compensation), in block order.
__host__ __device__ __forceinline__ double mass_combine(const double2* part) {
  double s = 0.0, comp = 0.0;
  for (int b = 0; b < kMassBlocks; ++b) {
    neumaier_add(s, comp, part[b].x);
    neumaier_add(s, comp, part[b].y);
  }
  return s + comp;
}

}  // namespace cn
