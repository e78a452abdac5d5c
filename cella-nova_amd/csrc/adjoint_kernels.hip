// Adjoint (reverse-mode) kernels of one DiffusionLenia step.  No counterpart in the reference,
// which has no differentiable GPU path; the PyTorch model (train/diff_lenia_org.py:27-53) is
// differentiated by autograd instead.  Forward step, per channel c and kernel k:
//   u_k = s_{src k} (*) K_k                      (circular convolution)
//   G_k = 2 exp(-(u_k - mu_k)^2 / (2 sigma_k^2)) - 1
//   A_c = exp(temp * sum_{k: dst k = c} w_k G_k)
//   Z_c = box3(A_c),  q_c = s_c / Z_c,  s'_c = A_c * box3(q_c)
// Given g = dL/ds', one adjoint step computes (with the forward intermediates recomputed):
//   T1 = box3(g A),  dL/ds (diffusion part) = T1 / Z,  dZ = -T1 q / Z
//   dL/dA = g box3(q) + box3(dZ),  gaff = dL/daff = temp A dL/dA
//   gu_k = w_k gaff_{dst k} (G_k + 1) (-(u_k - mu_k) / sigma_k^2)
//   dL/dmu_k = -sum gu_k,  dL/dsigma_k = -sum gu_k (u_k - mu_k) / sigma_k,  dL/dw_k = sum G_k gaff
//   dL/ds_c (convolution part) = sum_{k: src k = c} gu_k (correlated with) K_k
//                              = IFFT(sum_k FFT(gu_k) conj(K^_k))   (one inverse per source)
// Launches: k_diffusion_adj, k_growth_adj_row (growth adjoint + row R2C of gu_k),
// k_col_adj (forward column FFT of every gu_k, product with the conjugate spectrum, summed per
// source, one inverse column FFT), k_row_inv_add (row C2R + diffusion part), k_param_reduce.
#include <hip/hip_runtime.h>
#include "bufio.h"
#include "cfg.h"
#include "fft.h"
#include "launch.h"
#include "realfft.h"
#include "segio.h"
#include "types.h"

namespace cn {

constexpr int kAdjTW = 60, kAdjTH = 8, kAdjThreads = 256;   // output columns / rows per wave

// x(l-1) + x(l) + x(l+1) across the wavefront (DPP wave shifts; lanes 0 and 63 see a 0).
__device__ __forceinline__ float hsum3f(float x) {
  const float lo = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, false));
  const float hi = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, false));
  return lo + x + hi;
}

// ---------------------------------------------------------------------------------------
// Diffusion adjoint: inputs A (recomputed), state slot `sin`, gradient slot `gin`;
// outputs dsd = T1 / Z and gaff.  Same streaming layout as the forward diffusion: one
// wavefront per 60 x kAdjTH tile, lane l on column x0 - 2 + l, horizontal 3-sums by DPP wave
// shifts and vertical ones in rolling registers while walking down the rows.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kAdjThreads)
k_diffusion_adj(const CreatureDev* __restrict__ descs, int sin, int gin, int N, int islot) {
  constexpr int OW = kAdjTW, TH = kAdjTH;
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N) return;
  const int c = blockIdx.y;
  if (c >= d.C) return;
  const int tiles_x = (N + OW - 1) / OW, tiles_y = (N + TH - 1) / TH;
  const int nblk = gridDim.x, bx = (int)blockIdx.x;
  const int blk = (nblk % 8 == 0) ? (bx % 8) * (nblk / 8) + bx / 8 : bx;
  const int tile = blk * (kAdjThreads / 64) + (int)(threadIdx.x >> 6);
  if (tile >= tiles_x * tiles_y) return;  // a whole wavefront; the kernel has no barriers
  const int l = threadIdx.x & 63;
  const int y0 = (tile / tiles_x) * TH, x0 = (tile % tiles_x) * OW;
  const int mask = N - 1;
  const int xl = x0 - 2 + l;
  const int gx = xl & mask;
  const bool store = l >= 2 && l < 2 + OW && xl < N;
  const size_t plane = (size_t)N * N;
  gptr<const float> A = to_global((const float*)d.A) + ((size_t)islot * d.C + c) * plane + gx;
  gptr<const float> s = slot_ptr(d, sin) + c * plane + gx;
  gptr<const float> g = to_global((const float*)d.grad) + ((size_t)gin * d.C + c) * plane + gx;
  gptr<float> dsd = to_global(d.dsd) + c * plane + xl;
  gptr<float> gaff = to_global(d.gaff) + c * plane + xl;
  const float temp = d.temp;

  // All rows of the tile are loaded up front; step j handles q row yq = y0 - 1 + j and from
  // j = 2 on finishes gaff of row yq - 1.
  float ar[TH + 4], gr[TH + 4], sr[TH + 2];
#pragma unroll
  for (int j = 0; j < TH + 4; ++j) {
    const size_t o = (size_t)((y0 - 2 + j) & mask) * N;
    ar[j] = A[o];
    gr[j] = g[o];
  }
#pragma unroll
  for (int j = 0; j < TH + 2; ++j) sr[j] = s[(size_t)((y0 - 1 + j) & mask) * N];
  float hz0 = hsum3f(ar[0]), hz1 = hsum3f(ar[1]);                  // horizontal sums of A
  float ht0 = hsum3f(ar[0] * gr[0]), ht1 = hsum3f(ar[1] * gr[1]);  // ... of A g
  float hq0 = 0.f, hq1 = 0.f, hd0 = 0.f, hd1 = 0.f;        // ... of q and dZ (previous q rows)
  float a0 = 0.f, a1 = 0.f, g0 = 0.f, g1 = 0.f;            // A, g of the previous step rows
#pragma unroll
  for (int j = 0; j < TH + 2; ++j) {
    const int yq = y0 - 1 + j;
    const float a2 = ar[j + 2], g2 = gr[j + 2];
    const float sv = sr[j];
    const float hz2 = hsum3f(a2), ht2 = hsum3f(a2 * g2);
    const float z = hz0 + hz1 + hz2, t1 = ht0 + ht1 + ht2;
    // One reciprocal (within 1 ulp) instead of three IEEE divisions: the gradients' tolerance
    // is set by the FP32 forward, not by the last bit of 1/Z.
    const float rz = __builtin_amdgcn_rcpf(z);
    const float q = sv * rz;
    const float dz = -t1 * q * rz;
    if (store && j >= 1 && j <= TH && yq < N) dsd[(size_t)yq * N] = t1 * rz;
    const float hq2 = hsum3f(q), hd2 = hsum3f(dz);
    if (j >= 2) {
      const int yo = yq - 1;
      const float ga = g0 * (hq0 + hq1 + hq2) + (hd0 + hd1 + hd2);
      if (store && yo < N) gaff[(size_t)yo * N] = ga * temp * a0;
    }
    hz0 = hz1; hz1 = hz2;
    ht0 = ht1; ht1 = ht2;
    hq0 = hq1; hq1 = hq2;
    hd0 = hd1; hd1 = hd2;
    a0 = a1; a1 = a2;
    g0 = g1; g1 = g2;
  }
}

// ---------------------------------------------------------------------------------------
// Growth adjoint fused with the row R2C transform of gu_k; per-block sums for the
// parameter gradients.  Grid (N/ROWS, K, creatures).
// ---------------------------------------------------------------------------------------
template <int N, int ER>
__global__ void __launch_bounds__((RowCfg<N, ER>::THREADS))
k_growth_adj_row(const CreatureDev* __restrict__ descs, const float2* __restrict__ tw, int islot) {
  using CF = RowCfg<N, ER>;
  constexpr int L = CF::L, E = CF::E, T = CF::T, ROWS = CF::ROWS, NT = CF::THREADS;
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N) return;
  const int k = blockIdx.y;
  if (k >= d.K) return;
  const int lane = threadIdx.x / T, t = threadIdx.x % T;
  const int y = blockIdx.x * ROWS + lane;
  const float mu = d.mu[k], is2 = d.inv_s2[k], wk = d.w[k], sg = d.sigma[k];
  const int c = d.dst_of[k];
  gptr<const float2> u = c64((const void*)(d.u + (((size_t)islot * d.K + k) * N + y) * N));
  gptr<const float2> ga = c64((const void*)(d.gaff + ((size_t)c * N + y) * N));
  float2 v[E];
  float smu = 0.f, ssg = 0.f, sw = 0.f;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int n = in_pos<L, E, false>(t, i);
    const float2 uu = u[n], gg = ga[n];
    const float d0 = uu.x - mu, d1 = uu.y - mu;
    const float e0 = 2.0f * exp_fast(-0.5f * d0 * d0 * is2), e1 = 2.0f * exp_fast(-0.5f * d1 * d1 * is2);  // G + 1
    const float gu0 = wk * gg.x * e0 * (-d0 * is2), gu1 = wk * gg.y * e1 * (-d1 * is2);
    v[i] = make_float2(gu0, gu1);
    smu -= gu0 + gu1;
    ssg -= gu0 * d0 + gu1 * d1;
    sw += (e0 - 1.0f) * gg.x + (e1 - 1.0f) * gg.y;
  }
  ssg /= sg;  // dL/dsigma = -sum gu (u - mu) / sigma: one division per thread
  __shared__ float2 lds[LdsMap<L, ROWS, false>::size];
  __shared__ float red[3][NT];
  using TF = TwLds<L, E, false>;
  __shared__ float2 twl[TF::size + SplitTw<N>::size];  // LDS twiddles (see k_row_fwd)
  {
    TwFill<TF, SplitTw<N>, CF::THREADS> twf;
    twf.load(tw, threadIdx.x);
    twf.store(twl, threadIdx.x);
  }
  red[0][threadIdx.x] = smu;
  red[1][threadIdx.x] = ssg;
  red[2][threadIdx.x] = sw;
  __syncthreads();
  fft_run<L, E, ROWS, false, false, kTwN, 0, true>(v, lds, t, lane, twl);
  __syncthreads();  // orders the red[] writes above before the reduction
  for (int h = NT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 3) to_global(d.partial)[((size_t)blockIdx.x * d.K + k) * 3 + threadIdx.x] = red[threadIdx.x][0];
  r2c_finish<N, E, ROWS, true>(v, lds, t, lane, twl + TF::size,
                               c64(d.specg) + ((size_t)k * N + y) * L, true);
}

// ---------------------------------------------------------------------------------------
// Column adjoint: for one source channel and column tile, sum over the kernels reading that
// source of FFT_col(rowspec(gu_k)) * conj(K^_k), then one inverse column FFT into adjspec[c].
// ---------------------------------------------------------------------------------------
template <int N, int TC, int EC>
__global__ void __launch_bounds__((TC * ColCfg<N, EC>::T)) __attribute__((amdgpu_waves_per_eu(4)))
k_col_adj(const CreatureDev* __restrict__ descs, const float2* __restrict__ tw, int skip_seg) {
  constexpr int L = N, E = ColCfg<N, EC>::E, HALF = N / 2;
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N || (skip_seg && d.seg_g > 0)) return;  // skip_seg: the segmented pass covers those
  const int c = blockIdx.y;
  if (c >= d.C) return;
  const int kb = d.src_off[c], ke = d.src_off[c + 1];
  const int x0 = xcd_tile<(N / 2) / TC>(blockIdx.x) * TC;
  const int lane = threadIdx.x % TC;
  int t = threadIdx.x / TC;
  const int x = x0 + lane;
  const bool edge = (x0 == 0);
  __shared__ float2 lds[LdsMap<L, TC, true>::size];
  using TF = TwLds<L, E, false>;
  using TI = TwLds<L, E, true>;
  __shared__ float2 twl[TF::size + TI::size + 1];  // both plans' twiddles (see TwLds)
  {
    TwFill<TF, TI, TC * ColCfg<N, EC>::T> twf;
    twf.load(tw, threadIdx.x);
    twf.store(twl, threadIdx.x);
  }

  // acc += v * conj(K^_k) at this thread's positions (compact real copy when available)
  auto accumulate = [&](int k, const float2* v, float2* acc) {
    if (d.ksym[k]) {
      gptr<const float> kr = to_global(d.kreal) + (size_t)k * (HALF + 1) * HALF;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int y = out_pos<L, E, false>(t, i);
        const float r = kr[kreal_index(N, y <= HALF ? y : N - y, x)];
        acc[i] = make_float2(acc[i].x + v[i].x * r, acc[i].y + v[i].y * r);
      }
    } else {
      gptr<const float2> kh = c64(d.khalf) + (size_t)k * N * HALF;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int y = out_pos<L, E, false>(t, i);
        acc[i] = c_add(acc[i], c_mulc(v[i], kh[(size_t)y * HALF + x]));
      }
    }
  };
  float2 acc[E];
#pragma unroll
  for (int i = 0; i < E; ++i) acc[i] = make_float2(0.f, 0.f);
  for (int kk = kb; kk < ke; ++kk) {
    const int k = d.src_k[kk];
    t = launder(t);
    gptr<const float2> sp = c64((const void*)d.specg) + (size_t)k * N * HALF;
    gptr<const float2> kh = c64(d.khalf) + (size_t)k * N * HALF;
    float2 v[E];
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = sp[(size_t)in_pos<L, E, false>(t, i) * HALF + x];
    fft_run<L, E, TC, true, false, kTwN, 0, true>(v, lds, t, lane, twl);
    if (edge) {
      // Column 0 packs (DC column) + i (Nyquist column): split, multiply each by its
      // conjugate kernel column, re-pack.
      __syncthreads();
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < E; ++i) lds[LdsMap<L, TC, true>::at(out_pos<L, E, false>(t, i), 0)] = v[i];
      }
      __syncthreads();
      if (lane == 0) {
        gptr<const float2> kn = c64(d.knyq) + (size_t)k * N;
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const int y = out_pos<L, E, false>(t, i);
          const float2 F = v[i];
          const float2 G = lds[LdsMap<L, TC, true>::at((N - y) & (N - 1), 0)];
          const float2 a = make_float2(0.5f * (F.x + G.x), 0.5f * (F.y - G.y));
          const float2 dd = make_float2(F.x - G.x, F.y + G.y);
          const float2 b = make_float2(0.5f * dd.y, -0.5f * dd.x);
          const float2 pa = c_mulc(a, kh[(size_t)y * HALF]);
          const float2 pb = c_mulc(b, kn[y]);
          acc[i] = c_add(acc[i], make_float2(pa.x - pb.y, pa.y + pb.x));
        }
      } else {
        accumulate(k, v, acc);
      }
    } else {
      accumulate(k, v, acc);
    }
  }
  fft_run<L, E, TC, true, true, kTwN, 0, true>(acc, lds, t, lane, twl + TF::size);
  // Result layout of the column passes (res_index), as the forward column pass writes uhat.
  gptr<float2> out = c64(d.adjspec) + (size_t)c * N * HALF;
#pragma unroll
  for (int i = 0; i < E; ++i) out[res_index(N, out_pos<L, E, true>(t, i), x)] = acc[i];
}

// ---------------------------------------------------------------------------------------
// Segmented adjoint column pass, for creatures with the segmented forward pass (every kernel
// real and even, column profiles of support |y| <= seg_r): conj(Khat_h) = Khat_h, so the
// correlation with each profile is the forward convolution, and per source the sum over its
// kernels is taken on the segments' spectra before one inverse transform.  Jobs, segment rows,
// tables and the packed DC/Nyquist column as in k_col_segments; a source without kernels
// writes zeros, as k_col_adj does.
// ---------------------------------------------------------------------------------------
template <int N, int M, int E>
__global__ void __launch_bounds__((seg_tc(N) * (M / E)))
k_col_adj_segments(const CreatureDev* __restrict__ descs, const float2* __restrict__ tw) {
  constexpr int TC = seg_tc(N), HALF = N / 2, W = HALF + 1, NT = HALF / TC, MH = M / 2;
  constexpr uint32_t kPlaneBytes = (uint32_t)N * HALF * 8, kTabBytes = (uint32_t)(MH + 1) * W * 4;
  using LM = LdsMap<M, TC, true>;
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N) return;
  const int c = blockIdx.y;
  if (c >= d.C) return;
  const int G = d.seg_g, R = d.seg_r;
  const int nb = (int)gridDim.x, bx = (int)blockIdx.x, gmax = nb / NT;
  const int job = (nb % 8 == 0) ? (bx % 8) * (nb / 8) + bx / 8 : bx;
  const int tile = job / gmax, seg = job % gmax;
  if (seg >= G) return;
  const int kb = d.src_off[c], ke = d.src_off[c + 1];
  const int lane = (int)threadIdx.x % TC, x = tile * TC + lane;
  int t = (int)threadIdx.x / TC;
  const int o0 = seg * N / G, rows = (seg + 1) * N / G - o0, base = o0 - R;
  const bool edge = tile == 0;
  using TF = TwLds<M, E, false>;
  using TI = TwLds<M, E, true>;
  __shared__ float2 twl[TF::size + TI::size + 1];
  {
    TwFill<TF, TI, TC * (M / E)> twf;
    twf.load(tw, threadIdx.x);
    twf.store(twl, threadIdx.x);
  }
  __shared__ float2 lds[LM::size];
  float2 acc[E];
#pragma unroll
  for (int i = 0; i < E; ++i) acc[i] = make_float2(0.f, 0.f);
  using IO = SegIO<N, M, E>;
  for (int kk = kb; kk < ke; ++kk) {
    const int k = d.src_k[kk];
    t = launder(t);
    const BufRes kr = make_res(d.kseg + (size_t)k * (MH + 1) * W, kTabBytes);
    float2 v[E];
    float hk[E];
    IO::load((const char*)d.specg + (size_t)k * kPlaneBytes, base, t, x, v);
    IO::table(d.kseg + (size_t)k * (MH + 1) * W, t, x, hk);
    fft_run<M, E, TC, true, false, kTwN, 0, true>(v, lds, t, lane, twl);
    if (edge) {
      __syncthreads();
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < E; ++i) lds[LM::at(out_pos<M, E, false>(t, i), 0)] = v[i];
      }
      __syncthreads();
    }
    if (edge && lane == 0) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int m = out_pos<M, E, false>(t, i), mm = m <= MH ? m : M - m;
        const float hn = buf_load_f32(kr, (uint32_t)(mm * W + HALF) * 4u, 0);
        const float p = 0.5f * (hk[i] + hn), q = 0.5f * (hk[i] - hn);
        const float2 F = v[i], Gm = lds[LM::at((M - m) & (M - 1), 0)];
        acc[i] = make_float2(acc[i].x + F.x * p + Gm.x * q, acc[i].y + F.y * p - Gm.y * q);
      }
    } else {
#pragma unroll
      for (int i = 0; i < E; ++i) acc[i] = make_float2(acc[i].x + v[i].x * hk[i], acc[i].y + v[i].y * hk[i]);
    }
  }
  fft_run<M, E, TC, true, true, kTwN, 0, true>(acc, lds, t, lane, twl + TF::size);
  IO::store((char*)d.adjspec + (size_t)c * kPlaneBytes, acc, t, x, o0, R, rows);
}

constexpr int kReduceWaves = 4;

// One (kernel, parameter) pair idx = 3 k + j of creature d, on one full wavefront (lane).
__device__ __forceinline__ void param_reduce_pair(const CreatureDev& d, int idx, int lane, int nblk) {
  const int K = d.K;
  const int k = idx / 3, j = idx % 3;
  gptr<const float> part = to_global((const float*)d.partial);
  float s = 0.f, comp = 0.f;  // running sum and Kahan compensation (true sum = s - comp)
  for (int b = lane; b < nblk; b += 64) {
    const float yv = part[((size_t)b * K + k) * 3 + j] - comp;
    const float tt = s + yv;
    comp = (tt - s) - yv;
    s = tt;
  }
  for (int m = 32; m >= 1; m >>= 1) {
    const float s2 = __shfl_xor(s, m, 64), c2 = __shfl_xor(comp, m, 64);
    const float t = s + s2;  // TwoSum: s + s2 = t + e exactly
    const float bv = t - s;
    const float e = (s - (t - bv)) + (s2 - bv);
    comp = (comp + c2) - e;
    s = t;
  }
  if (lane == 0) {
    gptr<float> acc = to_global(d.gparam) + j * K + k;
    gptr<float> cmp = to_global(d.gparam) + 3 * K + j * K + k;
    const float yv = (s - comp) - *cmp;
    const float tt = *acc + yv;
    *cmp = (tt - *acc) - yv;
    *acc = tt;
  }
}

__global__ void __launch_bounds__(64 * kReduceWaves)
k_param_reduce(const CreatureDev* __restrict__ descs, int nblk) {
  const CreatureDev& d = descs[blockIdx.z];
  const int idx = (int)blockIdx.x * kReduceWaves + (int)(threadIdx.x >> 6);
  if (idx >= 3 * d.K) return;  // a whole wavefront; the kernel has no barriers
  param_reduce_pair(d, idx, (int)(threadIdx.x & 63), nblk);
}

// ---------------------------------------------------------------------------------------
// Row C2R of adjspec[c] (the convolution part of dL/ds) plus the diffusion part -> grad slot.
// ---------------------------------------------------------------------------------------
template <int N, int ER>
__global__ void __launch_bounds__((RowCfg<N, ER>::THREADS))
k_row_inv_add(const CreatureDev* __restrict__ descs, int gout, const float2* __restrict__ tw, int red_nblk) {
  using CF = RowCfg<N, ER>;
  constexpr int L = CF::L, E = CF::E, T = CF::T, ROWS = CF::ROWS;
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N) return;
  if (red_nblk > 0 && (int)blockIdx.y == (int)gridDim.y - 1) {
    // The last y slice reduces the growth adjoint's parameter partials (red_nblk per kernel),
    // which are final once the growth adjoint has run: one launch fewer per adjoint step.
    if constexpr (CF::THREADS % 64 == 0) {
      constexpr int W = CF::THREADS / 64;
      for (int idx = (int)blockIdx.x * W + (int)(threadIdx.x >> 6); idx < 3 * d.K; idx += (int)gridDim.x * W)
        param_reduce_pair(d, idx, (int)(threadIdx.x & 63), red_nblk);
    }
    return;
  }
  const int c = blockIdx.y;
  if (c >= d.C) return;
  const int lane = threadIdx.x / T, t = threadIdx.x % T;
  const int y = blockIdx.x * ROWS + lane;
  __shared__ float2 lds[LdsMap<L, ROWS, false>::size];
  using TI = TwLds<L, E, true>;
  __shared__ float2 twl[TI::size + SplitTw<N>::size];  // LDS twiddles (see k_row_fwd)
  {
    TwFill<TI, SplitTw<N>, CF::THREADS> twf;
    twf.load(tw, threadIdx.x);
    twf.store(twl, threadIdx.x);
  }
  __syncthreads();
  float2 v[E];
  c2r_start_ct<N, E, true>((const float2*)d.adjspec + (size_t)c * N * L, y, t, twl + TI::size, v);
  fft_run<L, E, ROWS, false, true, kTwN, 0, true>(v, lds, t, lane, twl);
  const size_t row = ((size_t)c * N + y) * N;
  gptr<const float2> dd = c64((const void*)(d.dsd + row));
  gptr<float2> o = c64((void*)(d.grad + (size_t)gout * d.C * N * N + row));
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int n = out_pos<L, E, true>(t, i);
    o[n] = c_add(dd[n], v[i]);
  }
}

// ---------------------------------------------------------------------------------------
// Parameter gradients: sum the per-block partials and add them to the accumulated gradients
// with Kahan compensation.  One wavefront per (kernel, parameter) pair: lane l sums partials
// l, l + 64, ... (Kahan), then the lanes combine in a fixed xor butterfly with TwoSum, so the
// result is the same on every run.  gparam = [mu | sigma | w] then the compensation terms.
// ---------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_zero_gparam(const CreatureDev* __restrict__ descs) {
  const CreatureDev& d = descs[blockIdx.z];
  for (int i = threadIdx.x; i < 6 * d.K; i += blockDim.x) d.gparam[i] = 0.f;
}

// ---------------------------------------------------------------------------------------
// MSE loss against the target: g = 2 (x - target) / (C N^2) into grad slot gout, per-block
// partial sums of (x - target)^2 into lossp.
// ---------------------------------------------------------------------------------------
constexpr int kLossThreads = 256, kLossBlocks = kLossPartials;

__global__ void __launch_bounds__(kLossThreads)
k_mse_grad(const CreatureDev* __restrict__ descs, int sfin, int gout) {
  const CreatureDev& d = descs[blockIdx.z];
  const size_t n = (size_t)d.C * d.N * d.N;
  gptr<const float> x = slot_ptr(d, sfin);
  gptr<const float> tg = to_global((const float*)d.target);
  gptr<float> g = to_global(d.grad) + (size_t)gout * n;
  const float scale = 2.0f / (float)n;
  float acc = 0.f;
  for (size_t i = (size_t)blockIdx.x * kLossThreads + threadIdx.x; i < n; i += (size_t)kLossBlocks * kLossThreads) {
    const float r = x[i] - tg[i];
    g[i] = scale * r;
    acc += r * r;
  }
  __shared__ float red[kLossThreads];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kLossThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) to_global(d.lossp)[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------
template <int N, int TC, int EC>
static hipError_t launch_col_adj(const StepLaunch& sl, hipStream_t st) {
  if constexpr (!col_tc_ok<N, TC, EC>()) {
    return hipErrorInvalidValue;
  } else {
    k_col_adj<N, TC, EC><<<dim3((N / 2) / TC, sl.max_c, sl.count), TC * ColCfg<N, EC>::T, 0, st>>>(
        sl.descs, sl.tw, sl.seg_g > 0 ? 1 : 0);
    return hipSuccess;
  }
}

template <int N, int EC>
static hipError_t launch_col_adj_tc(const StepLaunch& sl, int tc, hipStream_t st) {
  switch (tc) {
    case 2: return launch_col_adj<N, 2, EC>(sl, st);
    case 4: return launch_col_adj<N, 4, EC>(sl, st);
    case 8: return launch_col_adj<N, 8, EC>(sl, st);
    case 16: return launch_col_adj<N, 16, EC>(sl, st);
    default: return hipErrorInvalidValue;
  }
}

// red_nblk > 0: the inverse-row launch also reduces the parameter partials (an extra y slice);
// only for row geometries of whole wavefronts (RowCfg THREADS a multiple of 64).
template <int N, int ER>
static hipError_t launch_adj_rows(const StepLaunch& sl, int gout, hipStream_t st, bool growth, int red_nblk = 0) {
  using CF = RowCfg<N, ER>;
  if (growth) {
    k_growth_adj_row<N, ER><<<dim3(CF::BLOCKS, sl.max_k, sl.count), CF::THREADS, 0, st>>>(sl.descs, sl.tw, sl.islot);
  } else {
    const int gy = sl.max_c + (red_nblk > 0 ? 1 : 0);
    k_row_inv_add<N, ER><<<dim3(CF::BLOCKS, gy, sl.count), CF::THREADS, 0, st>>>(sl.descs, gout, sl.tw, red_nblk);
  }
  return hipSuccess;
}

// Parameter-gradient reduction of the growth adjoint's per-block partials (row geometry ER).
template <int N, int ER>
static void launch_param_reduce(const StepLaunch& sl, hipStream_t st) {
  k_param_reduce<<<dim3((3 * sl.max_k + kReduceWaves - 1) / kReduceWaves, 1, sl.count), 64 * kReduceWaves, 0, st>>>(
      sl.descs, RowCfg<N, ER>::BLOCKS);
}

// One adjoint step: (sl.adj_recompute) recompute the forward intermediates of state slot sl.sin
// (u, A) into intermediate slot sl.islot, then grad slot gin -> gout from those intermediates and
// parameter-gradient accumulation.
template <int N>
static hipError_t launch_adjoint_n(const StepLaunch& sl, int gin, int gout, hipStream_t st) {
  hipError_t e;
  if (sl.adj_recompute) {
    StepLaunch fw = sl;
    fw.write_u = true;
    fw.skip_diffusion = true;
    fw.ev[0] = fw.ev[1] = fw.ev[2] = fw.ev[3] = fw.ev[4] = nullptr;
    if ((e = launch_step(N, fw, st)) != hipSuccess) return e;
  }
  const int tiles_x = (N + kAdjTW - 1) / kAdjTW, tiles_y = (N + kAdjTH - 1) / kAdjTH;
  const int waves = kAdjThreads / 64;
  k_diffusion_adj<<<dim3((tiles_x * tiles_y + waves - 1) / waves, sl.max_c, sl.count), kAdjThreads, 0, st>>>(
      sl.descs, sl.sin, gin, N, sl.islot);
  auto rows = [&](bool growth, int red_nblk) {
    const int re = growth ? (sl.fwd_e ? sl.fwd_e : DefaultGeom<N>::FWD_E) : (sl.inv_e ? sl.inv_e : DefaultGeom<N>::INV_E);
    switch (re) {
      case 8: return launch_adj_rows<N, 8>(sl, gout, st, growth, red_nblk);
      case 16: return launch_adj_rows<N, 16>(sl, gout, st, growth, red_nblk);
      default: return hipErrorInvalidValue;
    }
  };
  // Partials of the growth adjoint: one per block of its row geometry (the plan's forward one).
  const int fe = sl.fwd_e ? sl.fwd_e : DefaultGeom<N>::FWD_E;
  const int pnblk = fe == 16 ? RowCfg<N, 16>::BLOCKS : RowCfg<N, 8>::BLOCKS;
  const int ie = sl.inv_e ? sl.inv_e : DefaultGeom<N>::INV_E;
  const bool fold = (ie == 16 ? RowCfg<N, 16>::THREADS : RowCfg<N, 8>::THREADS) % 64 == 0;
  if ((e = rows(true, 0)) != hipSuccess) return e;
  // The adjoint column kernel keeps a per-source accumulator beside the transform, so it runs
  // with 8 values per thread (16 would spill), on the plan's tile width or the widest valid one
  // below it, at most 4 columns (measured: wider tiles leave too few blocks, e.g. 192 blocks of
  // 1024 threads at 1024^2 C3K9 with 8-column tiles; 4-column tiles were 2-3 % faster per
  // rollout at 512^2 and 1024^2, 2-column tiles slower).
  int tc = sl.col_tc ? sl.col_tc : DefaultGeom<N>::TC;
  while (tc > 2 && tc * ColCfg<N, 8>::T > 1024) tc /= 2;
  if (tc > 4) tc = 4;
  if constexpr (seg_m(N) > 0) {
    if (sl.seg_g > 0) {
      // The forward pass's values per thread; with its per-source accumulator the kernel takes
      // more registers (133-147 VGPRs at 16 values, no spill, 0.8 % faster for the 512^2
      // batched search than 8 values at a 4-waves-per-SIMD register cap).
      constexpr int M = seg_m(N), SE = seg_e(N);
      k_col_adj_segments<N, M, SE><<<dim3((N / 2 / seg_tc(N)) * sl.seg_g, sl.max_c, sl.count), seg_tc(N) * (M / SE), 0,
                                      st>>>(sl.descs, sl.tw);
    }
  }
  if ((sl.seg_g == 0 || sl.seg_mixed) && (e = launch_col_adj_tc<N, 8>(sl, tc, st)) != hipSuccess) return e;
  // Parameter-gradient reduction over the partials of the growth adjoint, whose row geometry
  // (the plan's forward one) sets their block count; the inverse row geometry may differ.  It
  // rides on the inverse-row launch where that one runs whole wavefronts.
  if ((e = rows(false, fold ? pnblk : 0)) != hipSuccess) return e;
  if (!fold) {
    switch (fe) {
      case 8: launch_param_reduce<N, 8>(sl, st); break;
      case 16: launch_param_reduce<N, 16>(sl, st); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

hipError_t launch_adjoint(int N, const StepLaunch& sl, int gin, int gout, hipStream_t st) {
  switch (N) {
    case 16: return launch_adjoint_n<16>(sl, gin, gout, st);
    case 32: return launch_adjoint_n<32>(sl, gin, gout, st);
    case 64: return launch_adjoint_n<64>(sl, gin, gout, st);
    case 128: return launch_adjoint_n<128>(sl, gin, gout, st);
    case 256: return launch_adjoint_n<256>(sl, gin, gout, st);
    case 512: return launch_adjoint_n<512>(sl, gin, gout, st);
    case 1024: return launch_adjoint_n<1024>(sl, gin, gout, st);
    case 2048: return launch_adjoint_n<2048>(sl, gin, gout, st);
    default: return hipErrorInvalidValue;
  }
}

// Upper bound on the row-transform blocks per plane over all geometries (one row per block).
int row_blocks(int N) { return N; }

hipError_t launch_mse_grad(const CreatureDev* descs, int count, int sfin, int gout, hipStream_t st) {
  k_mse_grad<<<dim3(kLossBlocks, 1, count), kLossThreads, 0, st>>>(descs, sfin, gout);
  return hipGetLastError();
}

hipError_t launch_zero_gparam(const CreatureDev* descs, int count, hipStream_t st) {
  k_zero_gparam<<<dim3(1, 1, count), 256, 0, st>>>(descs);
  return hipGetLastError();
}

}  // namespace cn
