// Set-up of the segmented column pass (k_col_segments): column profiles of real, even kernel
// spectra, their support radius, and the per-segment tables.
//
// For kernel k and spectrum column kx, the column pass convolves the column of row spectra with
// the profile h(y, kx) = sum_ky Khat_h(ky, kx) e^{2 pi i ky y / N} (Khat_h pre-scaled by 1/N^2):
// a kernel of radius R in space has h(y, kx) = 0 for |y| > R, up to the f32 rounding of the
// stored spectrum.  A real spectrum that is even in ky gives a real profile that is even in y:
//   h(j) = Kr(0) + (-1)^j Kr(N/2) + 2 sum_{ky=1}^{N/2-1} Kr(ky) cos(2 pi ky j / N),  j = 0 .. N/2.
// The support radius is the smallest r whose tail energy sum_{|y| > r} h(y)^2, summed over all
// columns, is within (2^-22)^2 of the total (the profile's f32 rounding noise sits near 2^-25
// relative rms); the segment table of a kernel of support r is
//   H(m, kx) = (1/M) [h(0) + 2 sum_{j=1}^{r} h(j) cos(2 pi m j / M)],  m = 0 .. M/2,
// the M-point spectrum of the truncated profile with the inverse transform's 1/M folded in.
#include <hip/hip_runtime.h>
#include "cfg.h"
#include "launch.h"
#include "types.h"

namespace cn {

template <int N>
__global__ void __launch_bounds__(256)
k_seg_profile(const CreatureDev* __restrict__ desc, const int* __restrict__ ks, double* __restrict__ hbuf,
              double* __restrict__ etail) {
  constexpr int H = N / 2, W = H + 1, RC = kSegRcap;
  const int kx = blockIdx.x, kk = blockIdx.y, k = ks[kk];
  __shared__ double kr[W];
  __shared__ double ct[N];
  __shared__ double hs[W];  // w_j h(j)^2
  const CreatureDev& d = *desc;
  const float2* kh = (const float2*)d.khalf + (size_t)k * N * H;
  const float2* kn = (const float2*)d.knyq + (size_t)k * N;
  for (int ky = threadIdx.x; ky <= H; ky += blockDim.x)
    kr[ky] = kx < H ? (double)kh[(size_t)ky * H + kx].x : (double)kn[ky].x;
  for (int n = threadIdx.x; n < N; n += blockDim.x) ct[n] = cospi(2.0 * n / N);
  __syncthreads();
  for (int j = threadIdx.x; j <= H; j += blockDim.x) {
    double s = 0.0;
    for (int ky = 1; ky < H; ++ky) s = fma(kr[ky], ct[(ky * j) & (N - 1)], s);
    const double h = kr[0] + ((j & 1) ? -kr[H] : kr[H]) + 2.0 * s;
    hs[j] = (j == 0 || j == H ? 1.0 : 2.0) * h * h;
    if (j <= RC) hbuf[((size_t)kk * W + kx) * (RC + 1) + j] = h;
  }
  __syncthreads();
  for (int r1 = threadIdx.x; r1 <= RC + 1; r1 += blockDim.x) {
    double e = 0.0;
    for (int j = H; j >= r1; --j) e += hs[j];  // smallest terms first
    etail[((size_t)kk * W + kx) * (RC + 2) + r1] = e;
  }
}

// One block per (radius, kernel): strided partial sums over the columns, then a fixed-shape
// tree in LDS (the same order on every run).
__global__ void __launch_bounds__(256)
k_seg_tail_sum(const double* __restrict__ etail, int W, double* __restrict__ esum) {
  constexpr int RC = kSegRcap;
  const int r1 = blockIdx.x, kk = blockIdx.y;
  __shared__ double part[256];
  double e = 0.0;
  for (int kx = threadIdx.x; kx < W; kx += 256) e += etail[((size_t)kk * W + kx) * (RC + 2) + r1];
  part[threadIdx.x] = e;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) esum[kk * (RC + 2) + r1] = part[0];
}

template <int M>
__global__ void __launch_bounds__(256)
k_seg_table(const double* __restrict__ hbuf, const int* __restrict__ ks, const int* __restrict__ rs, int W,
            float* __restrict__ kseg) {
  constexpr int RC = kSegRcap;
  const int kx = blockIdx.x, kk = blockIdx.y, k = ks[kk], r = rs[kk];
  const double* h = hbuf + ((size_t)kk * W + kx) * (RC + 1);
  for (int m = threadIdx.x; m <= M / 2; m += blockDim.x) {
    double s = 0.0;
    for (int j = r; j >= 1; --j) s = fma(h[j], cospi(2.0 * ((m * j) % M) / M), s);
    kseg[((size_t)k * (M / 2 + 1) + m) * W + kx] = (float)((h[0] + 2.0 * s) / M);
  }
}

hipError_t launch_seg_profile(int N, const CreatureDev* desc, const int* ks, int nk, double* hbuf, double* etail,
                              hipStream_t st) {
  const dim3 g(N / 2 + 1, nk);
  switch (N) {
    case 2048: k_seg_profile<2048><<<g, 256, 0, st>>>(desc, ks, hbuf, etail); break;
    case 1024: k_seg_profile<1024><<<g, 256, 0, st>>>(desc, ks, hbuf, etail); break;
    case 512: k_seg_profile<512><<<g, 256, 0, st>>>(desc, ks, hbuf, etail); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_seg_tail_sum(int N, const double* etail, int nk, double* esum, hipStream_t st) {
  k_seg_tail_sum<<<dim3(kSegRcap + 2, nk), 256, 0, st>>>(etail, N / 2 + 1, esum);
  return hipGetLastError();
}

hipError_t launch_seg_table(int N, const double* hbuf, const int* ks, const int* rs, int nk, float* kseg,
                            hipStream_t st) {
  const dim3 g(N / 2 + 1, nk);
  switch (seg_m(N)) {
    case 256: k_seg_table<256><<<g, 256, 0, st>>>(hbuf, ks, rs, N / 2 + 1, kseg); break;
    case 128: k_seg_table<128><<<g, 256, 0, st>>>(hbuf, ks, rs, N / 2 + 1, kseg); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace cn
