// Shared host/device layout of one creature (one DiffusionLenia world) on the device.
//
// HBM layout (all planes row-major, channel/kernel-major):
//   state   [nslots][C][N][N] f32  world-state slots: 0/1 ping-pong current state (the
//                                  reference's channel/new_channel), the rest are rollout
//                                  checkpoints and recompute buffers.  Kernels take slot indices.
//   spec    [C][N][N/2]  c64   row-R2C spectra of a state (row-major)
//   khalf   [K][N][N/2]  c64   Hermitian part of each kernel spectrum, pre-scaled by 1/N^2,
//                              columns kx = 0 .. N/2-1
//   knyq    [K][N]       c64   the kx = N/2 (Nyquist) column of the same spectra
//   kreal   [K][N/8][N/2+1][4] f32 compact copy for kernels whose spectrum is real and even in
//                              ky (rows ky = 0 .. N/2; row N-ky equals row ky), in tiles of 4
//                              columns (kreal_index) so that column tiles read contiguous runs;
//                              ksym[k] = 1
//   kseg    [K][M/2+1][N/2+1] f32 segment tables of the segmented column pass (seg_g > 0):
//                              H_k(m, kx) = (1/M) DFT_M of kernel k's column profile truncated
//                              to |y| <= its support, rows m = 0 .. M/2 (real, even in m)
//   uhat    [K][N][N/2]  c64   row spectra of every convolution result (after the inverse
//                              column pass), row-major (res_index; column tiles ct_index when
//                              kColTiled); slot kx = 0 packs (DC, Nyquist) as (re, im)
//   A       [islots][C][N][N] f32 affinity exp(temp * sum_k w_k G_k) per destination channel;
//                              slot 0 for stepping, one slot per step of a rollout segment
//   u       [islots][K][N][N] f32 convolution results (readback and adjoint only)
//   render  [N][N]       u32   packed 0x00RRGGBB frame of channels 0..2 (allocated on the first
//                              iterate_and_render; the diffusion pass writes bytes 2, 1, 0)
// Adjoint buffers (allocated on first rollout):
//   grad    [2][C][N][N] f32   ping-pong dL/dstate
//   gaff    [C][N][N]    f32   dL/d(pre-exp affinity)
//   dsd     [C][N][N]    f32   diffusion part of dL/dstate
//   specg   [K][N][N/2]  c64   row spectra of dL/du_k
//   adjspec [C][N][N/2]  c64   per-source sums of the column adjoint (res_index layout), input of
//                              the inverse-row adjoint
//   partial [nblk][K][3] f32   per-block sums for dL/dmu, dL/dsigma, dL/dw
//   gparam  [2][3][K]    f32   accumulated (mu, sigma, w) gradients and their Kahan terms
//   target  [C][N][N]    f32   loss target
//   lossp   [nblk]       f32   per-block loss partial sums
//   adam    [6K + 2]     f32   Adam moments of (mu, sigma, w) and the bias-correction products
//                              (allocated on the first cn_adam_step)
#pragma once
#include <stdint.h>

namespace cn {

constexpr int kMaxChannels = 16;   // reference growth shader limit: array<f32,16>
constexpr int kMaxKernels = 256;
constexpr unsigned kMaxCreatures = 65535;  // per engine: the grid z dimension of a launch
constexpr int kTwN = 4096;         // twiddle table length; supports N <= 4096

// Twiddle table layout: kTwN entries tw[i] = exp(-2 pi i i / kTwN), then one block per Stockham
// pass shape (span NS, radix R), NS = 2 .. kTwN/R, R = 2, 4, 8, 16: entry (m - 1) * NS + k of
// the block is W_{NS R}^{m k}, m = 1..R-1, k = 0..NS-1 (the same FP64-rounded value as
// tw[m k kTwN / (NS R)]).  With k fastest, the lanes of a wavefront (consecutive k) read
// consecutive entries.
constexpr int tw_pass_offset(int NS, int R) {
  int off = kTwN;
  for (int r = 2; r <= 16; r *= 2)
    for (int ns = 2; ns * r <= kTwN; ns *= 2) {
      if (r == R && ns == NS) return off;
      off += (r - 1) * ns;
    }
  return -1;
}
constexpr int tw_table_size() {
  int n = kTwN;
  for (int r = 2; r <= 16; r *= 2)
    for (int ns = 2; ns * r <= kTwN; ns *= 2) n += (r - 1) * ns;
  return n;
}

struct CreatureDev {
  float* state;
  int nslots;
  void* spec;          // float2*
  const void* khalf;   // float2*
  const void* knyq;    // float2*
  const float* kreal;
  void* uhat;          // float2*
  float* A;
  float* u;
  unsigned int* render;
  float* grad;
  float* gaff;
  float* dsd;
  void* specg;         // float2*
  void* adjspec;       // float2*: the adjoint column pass's output (its own plane, so that a
                       // segment's recompute may run beside another segment's adjoint)
  float* partial;
  float* gparam;
  float* target;
  float* lossp;
  float* adam;         // [3K] first moments, [3K] second moments, beta1^t, beta2^t (optimiser)
  int param_ofs;       // offset of this creature's 3K parameters in a batch-wide array
  int N, C, K;
  float temp;
  // kernels grouped by source channel (CSR): src_k[src_off[c] .. src_off[c+1]) in increasing k
  int src_off[kMaxChannels + 1];
  int src_k[kMaxKernels];
  // kernels grouped by destination channel (CSR), increasing k inside a group
  int dst_off[kMaxChannels + 1];
  int dst_k[kMaxKernels];
  int src_of[kMaxKernels];     // src[k]
  int dst_of[kMaxKernels];     // dst[k]
  float mu[kMaxKernels];
  float sigma[kMaxKernels];
  float inv_s2[kMaxKernels];   // 1 / sigma^2
  float w[kMaxKernels];
  int ksym[kMaxKernels];       // 1: kernel k is read from kreal (real, even in ky)
  // Segmented column pass: seg_g segments of M rows (seg_m(N)) overlapping by 2 seg_r rows,
  // for creatures whose every kernel has a real, even spectrum and column profiles within
  // |y| <= seg_r; seg_g = 0: full-length column transforms.
  const float* kseg;
  int seg_g, seg_r;
};

// Offset (in complex elements) of row y, column x of a column-tiled half-spectrum plane
// (N rows, N/2 columns): the plane is N/8 tiles of 4 columns, each tile N rows of 4 values.
#ifdef __HIPCC__
__host__ __device__
#endif
inline size_t ct_index(int N, int y, int x) {
  return ((size_t)(x >> 2) * N + y) * 4 + (x & 3);
}

// Layout of the column pass's result planes (uhat, and the adjoint column pass's output in
// spec): column tiles (ct_index) when kColTiled, else row-major rows like the other planes.
// Row-major is the default: column tiles let the column pass write whole 128-byte lines (its
// launch 84.5 -> 79.0 us at 2048^2 C3K9) but give the inverse-row pass 16 partial lines per
// load instruction (52.3 -> 61.0 us), so the step is faster row-major (0.1797 against
// 0.1814 ms per step, same run).
#ifndef CN_COL_TILED
#define CN_COL_TILED 0
#endif
constexpr bool kColTiled = CN_COL_TILED != 0;

#ifdef __HIPCC__
__host__ __device__
#endif
inline size_t res_index(int N, int y, int x) {
  return kColTiled ? ct_index(N, y, x) : (size_t)y * (N / 2) + x;
}

// Offset of (ky <= N/2, kx < N/2) in one kernel's kreal block: 4-column tiles, rows inside.
#ifdef __HIPCC__
__host__ __device__
#endif
inline size_t kreal_index(int N, int ky, int kx) {
  return ((size_t)(kx >> 2) * (N / 2 + 1) + ky) * 4 + (kx & 3);
}

#ifdef __HIPCC__
// Global-memory views of the descriptor's device pointers.  Accesses through a generic pointer
// loaded from the descriptor compile to FLAT instructions, which count on the LDS/scalar-memory
// counter as well as on the vector-memory one: a wait for an LDS exchange or an LDS twiddle
// then also waits for every device-memory load and store in flight (prefetched spectra, the
// previous result's stores).  Address space 1 makes them global loads and stores, counted on
// the vector-memory counter only.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = T __attribute__((address_space(1)))*;
#else
template <class T>
using gptr = T*;  // host pass: kernel bodies are parsed there but never generated
#endif
template <class T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
  return (gptr<T>)p;
}
__device__ __forceinline__ gptr<float> slot_ptr(const CreatureDev& d, int s) {
  return to_global(d.state + (size_t)s * d.C * d.N * d.N);
}
// Complex (float2) views of the spectrum buffers.
__device__ __forceinline__ gptr<float2> c64(void* p) { return to_global(reinterpret_cast<float2*>(p)); }
__device__ __forceinline__ gptr<const float2> c64(const void* p) {
  return to_global(reinterpret_cast<const float2*>(p));
}
#endif

}  // namespace cn
