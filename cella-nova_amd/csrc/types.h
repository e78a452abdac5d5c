// Shared host/device layout of one creature (one DiffusionLenia world) on the device.
//
// HBM layout (all planes row-major, channel/kernel-major, 256-byte aligned bases):
//   state[2]  [C][N][N]    f32   ping-pong world state (the reference's channel/new_channel)
//   spec      [C][N][N/2]  c64   row-R2C spectra of the state, then column-transformed in place
//                                by nobody (the fused column pass writes uhat instead)
//   khalf     [K][N][N/2]  c64   Hermitian part of each kernel spectrum, pre-scaled by 1/N^2,
//                                columns kx = 0 .. N/2-1
//   knyq      [K][N]       c64   the kx = N/2 (Nyquist) column of the same spectra
//   uhat      [K][N][N/2]  c64   row spectra of every convolution result (after the inverse
//                                column pass); slot kx = 0 packs (DC, Nyquist) as (re, im)
//   A         [C][N][N]    f32   affinity exp(temp * sum_k w_k G_k) per destination channel
//   u         [K][N][N]    f32   optional: convolution results (only for readback/adjoints)
#pragma once
#include <stdint.h>

namespace cn {

constexpr int kMaxChannels = 16;   // reference growth shader limit: array<f32,16>
constexpr int kMaxKernels = 256;
constexpr int kTwN = 4096;         // twiddle table length; supports N <= 4096

struct CreatureDev {
  float* state[2];
  void* spec;          // float2*
  const void* khalf;   // float2*
  const void* knyq;    // float2*
  void* uhat;          // float2*
  float* A;
  float* u;            // may be null
  int N, C, K;
  float temp;
  // kernels grouped by source channel (CSR): src_k[src_off[c] .. src_off[c+1]) in increasing k
  int src_off[kMaxChannels + 1];
  int src_k[kMaxKernels];
  // kernels grouped by destination channel (CSR), increasing k inside a group
  int dst_off[kMaxChannels + 1];
  int dst_k[kMaxKernels];
  float mu[kMaxKernels];
  float inv_s2[kMaxKernels];   // 1 / sigma^2
  float w[kMaxKernels];
};

}  // namespace cn
