// Device-side optimiser step and batch-wide parameter copies for creature searches.
//
// The reference optimises its creatures' growth parameters with torch.optim.Adam and clamps them
// to the box of LeniaParams._sanitize (train/utils/leniaparams.py:351-360): mu in [0, 2],
// sigma >= a floor, weights >= 0.  Here the update runs on the device from the gradients the
// last backward left in each creature's descriptor and writes the new parameters into the
// descriptor the step kernels read, so an optimisation loop needs no host round trip.
#include <hip/hip_runtime.h>
#include "launch.h"
#include "types.h"

namespace cn {

// One Adam step per creature (blockIdx.z): m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps), then the sanitize box.  FP32 with
// no contraction, so that a host restatement in float32 gives the same bits.
__global__ void __launch_bounds__(256) k_adam(CreatureDev* __restrict__ descs, AdamParams p) {
#pragma clang fp contract(off)
  CreatureDev& d = descs[blockIdx.z];
  const int K = d.K;
  gptr<float> st = to_global(d.adam);
  gptr<const float> g = to_global((const float*)d.gparam);
  const float b1t = st[6 * K] * p.beta1, b2t = st[6 * K + 1] * p.beta2;
  for (int j = threadIdx.x; j < 3 * K; j += blockDim.x) {
    const float gj = g[j];
    const float m = p.beta1 * st[j] + (1.0f - p.beta1) * gj;
    const float v = p.beta2 * st[3 * K + j] + (1.0f - p.beta2) * (gj * gj);
    st[j] = m;
    st[3 * K + j] = v;
    const float mh = m / (1.0f - b1t), vh = v / (1.0f - b2t);
    const int q = j / K, k = j - q * K;
    float* par = q == 0 ? d.mu : (q == 1 ? d.sigma : d.w);
    float x = par[k] - p.lr * mh / (sqrtf(vh) + p.eps);
    if (q == 0) x = fminf(fmaxf(x, p.mu_lo), p.mu_hi);
    else if (q == 1) x = fmaxf(x, p.sigma_lo);
    else x = fmaxf(x, p.w_lo);
    par[k] = x;
    if (q == 1) d.inv_s2[k] = 1.0f / (x * x);
  }
  __syncthreads();  // every thread has read the products before they advance
  if (threadIdx.x == 0) {
    st[6 * K] = b1t;
    st[6 * K + 1] = b2t;
  }
}

// Moments to zero and products to one (the state before the first step).
__global__ void __launch_bounds__(256) k_adam_reset(const CreatureDev* __restrict__ descs) {
  const CreatureDev& d = descs[blockIdx.z];
  gptr<float> st = to_global(d.adam);
  for (int j = threadIdx.x; j < 6 * d.K + 2; j += blockDim.x) st[j] = j < 6 * d.K ? 0.0f : 1.0f;
}

// The 3K parameter gradients (mu, sigma, w) of every creature into out[param_ofs ...].
__global__ void __launch_bounds__(256) k_gather_grads(const CreatureDev* __restrict__ descs, float* __restrict__ out) {
  const CreatureDev& d = descs[blockIdx.z];
  gptr<const float> g = to_global((const float*)d.gparam);
  for (int j = threadIdx.x; j < 3 * d.K; j += blockDim.x) out[d.param_ofs + j] = g[j];
}

hipError_t launch_adam(CreatureDev* descs, int count, const AdamParams& p, hipStream_t st) {
  k_adam<<<dim3(1, 1, count), 256, 0, st>>>(descs, p);
  return hipGetLastError();
}

hipError_t launch_adam_reset(const CreatureDev* descs, int count, hipStream_t st) {
  k_adam_reset<<<dim3(1, 1, count), 256, 0, st>>>(descs);
  return hipGetLastError();
}

hipError_t launch_gather_grads(const CreatureDev* descs, int count, float* out, hipStream_t st) {
  k_gather_grads<<<dim3(1, 1, count), 256, 0, st>>>(descs, out);
  return hipGetLastError();
}

}  // namespace cn
