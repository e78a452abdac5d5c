// Host-side launch interface of the device kernels (shared by the engine and the batch path).
#pragma once
#include <hip/hip_runtime.h>
#include "cfg.h"
#include "prng.h"
#include "types.h"

namespace cn {

struct StepLaunch {
  const CreatureDev* descs = nullptr;  // device array of `count` descriptors, all with the same N
  int count = 0;
  int sin = 0, sout = 1;               // state slot read / written
  int max_c = 0;                       // max channels over the descriptors
  int max_k = 0;                       // max kernels over the descriptors
  const float2* tw = nullptr;          // kTwN-entry twiddle table
  bool write_u = false;                // also write the real convolution results to desc.u
  bool skip_diffusion = false;         // stop after the affinity (conv readback)
  bool render = false;                 // diffusion also packs channels 0..2 into desc.render
  // Slot of the step intermediates written (inverse-row pass) and read (diffusion, adjoint):
  // the affinity at desc.A + islot*C*N*N and the convolution results at desc.u + islot*K*N*N.
  int islot = 0;
  bool adj_recompute = true;           // adjoint: recompute u and A of sin into islot first
  // Launch geometry (0: built-in default): values per thread of the forward (R2C) and
  // inverse (C2R) row FFTs and of the column FFTs, column tile width, diffusion tile height.
  int fwd_e = 0, inv_e = 0, col_e = 0, col_tc = 0, diff_th = 0;
  bool col_split = false;              // column pass per (tile, kernel) instead of (tile, source)
  // > 0: the segmented column pass over seg_g segments per column (the largest desc.seg_g of
  // the launch; every descriptor has seg_g > 0), instead of full-length column transforms.
  int seg_g = 0;
  bool seg_mixed = false;              // some descriptors have seg_g = 0: full-length pass for them
  // Optional per-kernel timing events: ev[0] before the row pass, ev[1] after it, ev[2] after
  // the column pass, ev[3] after the inverse-row + growth pass, ev[4] after the diffusion.
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

hipError_t launch_step(int N, const StepLaunch& sl, hipStream_t st);
// One adjoint step of state slot sl.sin: gradient slot gin (dL/dstate after the step) ->
// gout (dL/dstate before it); parameter gradients accumulate into desc.gparam.
hipError_t launch_adjoint(int N, const StepLaunch& sl, int gin, int gout, hipStream_t st);
// dL/dx = 2 (x - target) / numel into gradient slot gout, loss partials into desc.lossp.
hipError_t launch_mse_grad(const CreatureDev* descs, int count, int sfin, int gout, hipStream_t st);
hipError_t launch_zero_gparam(const CreatureDev* descs, int count, hipStream_t st);
int row_blocks(int N);  // row-transform blocks per plane (parameter-gradient partial count)
constexpr int kLossPartials = 256;

// Device creature generator: ring kernels of `desc` (K = desc->K, parameters in `rings`, whose
// `tot` the launch fills) -> khalf, knyq, kreal; uses desc->uhat as scratch.
hipError_t launch_generate(int N, const CreatureDev* desc, RingParams* rings, int K, const float2* tw,
                           hipStream_t st);
// Seed patch of every channel into state slot `slot` (`cells` = C N^2).
hipError_t launch_seed(const CreatureDev* desc, int slot, size_t cells, uint32_t side, uint32_t origin,
                       uint64_t seed, uint64_t creature, hipStream_t st);

// Device-side Adam on every creature's (mu, sigma, w) from its last backward's gradients, then
// the box mu in [mu_lo, mu_hi], sigma >= sigma_lo, w >= w_lo (writes the device descriptors).
struct AdamParams {
  float lr, beta1, beta2, eps;
  float mu_lo, mu_hi, sigma_lo, w_lo;
};
hipError_t launch_adam(CreatureDev* descs, int count, const AdamParams& p, hipStream_t st);
hipError_t launch_adam_reset(const CreatureDev* descs, int count, hipStream_t st);
// The 3K parameter gradients of every creature into out[desc.param_ofs ...].
hipError_t launch_gather_grads(const CreatureDev* descs, int count, float* out, hipStream_t st);

hipError_t launch_mass(const CreatureDev* descs, int count, int slot, int max_c, double2* partial, int nblk,
                       hipStream_t st);

// Segmented column pass set-up, for the nk kernels ks[0..nk) of creature `desc` (grid size N,
// real, even spectra):
//   launch_seg_profile: column profiles h(j, kx), j = 0 .. N/2 (FP64), of every column
//     kx = 0 .. N/2: hbuf[kk][kx][j] for j <= kSegRcap, and the tail energies
//     etail[kk][kx][r + 1] = sum_{j > r} w_j h(j, kx)^2 for r = -1 .. kSegRcap;
//   launch_seg_tail_sum: esum[kk][r + 1] = sum over kx of etail (fixed order);
//   launch_seg_table: kseg rows of kernel ks[kk] from hbuf, truncated at radius rs[kk].
hipError_t launch_seg_profile(int N, const CreatureDev* desc, const int* ks, int nk, double* hbuf, double* etail,
                              hipStream_t st);
hipError_t launch_seg_tail_sum(int N, const double* etail, int nk, double* esum, hipStream_t st);
hipError_t launch_seg_table(int N, const double* hbuf, const int* ks, const int* rs, int nk, float* kseg,
                            hipStream_t st);

// Per-channel byte sums of a packed 0x00RRGGBB frame of `cells` pixels (a multiple of 4):
// kStatsBlocks blocks each write [sum r, g, b, sum r^2, g^2, b^2] to partial[6 * block ...].
constexpr int kStatsThreads = 256;
constexpr int kStatsBlocks = 256;
hipError_t launch_render_stats(const uint32_t* frame, size_t cells, unsigned long long* partial, hipStream_t st);

}  // namespace cn
