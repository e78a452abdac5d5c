// Device creature generator: Lenia ring kernels evaluated on the grid in FP64, normalised in
// FP64 and transformed to the engine's spectrum storage with the same radix-R FFT building
// blocks as the step (the reference builds kernels on the host and FFTs them with numpy,
// train/generate_kernel_json.py:106-145, then uploads K*N*N complex64 through set_kernel).
//
//   k_gen_norm   one block per kernel: FP64 sum of the kernel over its support box
//   k_gen_rows   kernel rows evaluated in registers (x[2n] + i x[2n+1] packing) -> R2C half
//                spectra rows, written to the uhat buffer (free outside a step)
//   k_gen_cols   forward column FFT of those rows; splits the packed DC/Nyquist column and
//                writes khalf, knyq and the compact real copy kreal, all scaled by 1/N^2
//   k_gen_seed   seed patch of every channel from the counter PRNG (host-identical values)
#include <hip/hip_runtime.h>
#include "cfg.h"
#include "fft.h"
#include "prng.h"
#include "realfft.h"
#include "types.h"

namespace cn {

constexpr int kGenNormThreads = 256;

__global__ void __launch_bounds__(kGenNormThreads) k_gen_norm(RingParams* __restrict__ rings, int N) {
  const RingParams rp = rings[blockIdx.x];
  // Support: r < 1 implies |dx|, |dy| < R.  Offsets live in [-N/2, N/2 - 1].
  const int h = (int)ceil(rp.R);
  const int lo = h < N / 2 ? -h : -N / 2;
  const int hi = h < N / 2 - 1 ? h : N / 2 - 1;
  const int w = hi - lo + 1;
  double s = 0.0;
  for (int idx = threadIdx.x; idx < w * w; idx += kGenNormThreads)
    s += ring_value(rp, (double)(lo + idx % w), (double)(lo + idx / w));
  __shared__ double red[kGenNormThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int m = kGenNormThreads / 2; m > 0; m >>= 1) {
    if ((int)threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) rings[blockIdx.x].tot = red[0];
}

template <int N, int ER>
__global__ void __launch_bounds__((RowCfg<N, ER>::THREADS))
k_gen_rows(const CreatureDev* __restrict__ desc, const RingParams* __restrict__ rings, const float2* __restrict__ tw) {
  using CF = RowCfg<N, ER>;
  constexpr int L = CF::L, E = CF::E, T = CF::T, ROWS = CF::ROWS;
  const CreatureDev& d = *desc;
  const int nrows = d.K * N;
  const int row0 = blockIdx.x * ROWS;
  if (row0 >= nrows) return;
  const int lane = threadIdx.x / T, t = threadIdx.x % T;
  int row = row0 + lane;
  const bool valid = row < nrows;
  if (!valid) row = nrows - 1;
  const int k = row / N, y = row % N;
  const RingParams rp = rings[k];
  const double dy = wrap_offset((uint32_t)y, N);
  float2 v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int p = in_pos<L, E, false>(t, i);
    double a = ring_value(rp, wrap_offset(2 * p, N), dy);
    double b = ring_value(rp, wrap_offset(2 * p + 1, N), dy);
    if (rp.tot > 0.0) {
      a /= rp.tot;
      b /= rp.tot;
    }
    v[i] = make_float2((float)a, (float)b);
  }
  __shared__ float2 lds[LdsMap<L, ROWS, false>::size];
  using TF = TwLds<L, E, false>;
  __shared__ float2 twl[TF::size + SplitTw<N>::size];  // LDS twiddles (see k_row_fwd)
  {
    TwFill<TF, SplitTw<N>, CF::THREADS> twf;
    twf.load(tw, threadIdx.x);
    twf.store(twl, threadIdx.x);
  }
  __syncthreads();
  fft_run<L, E, ROWS, false, false, kTwN, 0, true>(v, lds, t, lane, twl);
  r2c_finish<N, E, ROWS, true>(v, lds, t, lane, twl + TF::size, c64(d.uhat) + (size_t)row * L,
                               valid);
}

template <int N, int TC, int EC>
__global__ void __launch_bounds__((TC * ColCfg<N, EC>::T))
k_gen_cols(const CreatureDev* __restrict__ desc, const float2* __restrict__ tw) {
  constexpr int L = N, E = ColCfg<N, EC>::E, HALF = N / 2;
  const CreatureDev& d = *desc;
  const int k = blockIdx.y;
  const int x0 = blockIdx.x * TC;
  const int lane = threadIdx.x % TC;
  const int t = threadIdx.x / TC;
  const int x = x0 + lane;
  gptr<const float2> src = c64((const void*)d.uhat) + (size_t)k * N * HALF;
  float2 v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = src[(size_t)in_pos<L, E, false>(t, i) * HALF + x];
  __shared__ float2 lds[LdsMap<L, TC, true>::size];
  using TF = TwLds<L, E, false>;
  __shared__ float2 twl[TF::size + 1];  // forward-plan twiddles (see TwLds)
  {
    TwFill<TF, TwNone, TC * ColCfg<N, EC>::T> twf;
    twf.load(tw, threadIdx.x);
    twf.store(twl, threadIdx.x);
  }
  fft_run<L, E, TC, true, false, kTwN, 0, true>(v, lds, t, lane, twl);

  const float sc = 1.0f / ((float)N * (float)N);  // exact: N is a power of two
  gptr<float2> kh = c64(const_cast<void*>(d.khalf)) + (size_t)k * N * HALF;
  gptr<float2> kn = c64(const_cast<void*>(d.knyq)) + (size_t)k * N;
  gptr<float> kr = to_global(const_cast<float*>(d.kreal)) + (size_t)k * (HALF + 1) * HALF;
  if (x0 == 0) {
    // Column 0 holds two real columns packed as (DC column) + i (Nyquist column): split its
    // transform F into A = (F + conj F(-y)) / 2 and B = (F - conj F(-y)) / (2i).
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < E; ++i) lds[LdsMap<L, TC, true>::at(out_pos<L, E, false>(t, i), 0)] = v[i];
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int y = out_pos<L, E, false>(t, i);
        const float2 F = v[i];
        const float2 G = lds[LdsMap<L, TC, true>::at((N - y) & (N - 1), 0)];
        const float2 dd = make_float2(F.x - G.x, F.y + G.y);
        kn[y] = make_float2(0.5f * dd.y * sc, -0.5f * dd.x * sc);
        v[i] = make_float2(0.5f * (F.x + G.x), 0.5f * (F.y - G.y));
      }
    }
  }
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int y = out_pos<L, E, false>(t, i);
    const float2 val = make_float2(v[i].x * sc, v[i].y * sc);
    kh[(size_t)y * HALF + x] = val;
    if (y <= HALF) kr[kreal_index(N, y, x)] = val.x;
  }
}

__global__ void k_gen_seed(const CreatureDev* __restrict__ desc, int slot, uint32_t side, uint32_t origin,
                           uint64_t seed, uint64_t creature) {
  const CreatureDev& d = *desc;
  const size_t P = (size_t)d.N * d.N;
  const size_t total = (size_t)d.C * P;
  gptr<float> s = slot_ptr(d, slot);
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const uint32_t c = (uint32_t)(idx / P);
    const uint32_t y = (uint32_t)((idx % P) / d.N), x = (uint32_t)(idx % d.N);
    float v = 0.0f;
    if (y >= origin && y < origin + side && x >= origin && x < origin + side)
      v = seed_value(seed, creature, c, y - origin, x - origin, side);
    s[idx] = v;
  }
}

template <int N>
static hipError_t launch_generate_n(const CreatureDev* desc, RingParams* rings, int K, const float2* tw,
                                    hipStream_t st) {
  using CF = RowCfg<N, DefaultGeom<N>::FWD_E>;
  constexpr int TC = DefaultGeom<N>::TC, EC = DefaultGeom<N>::COL_E;
  static_assert(col_tc_ok<N, TC, EC>(), "default column geometry must be valid");
  k_gen_norm<<<K, kGenNormThreads, 0, st>>>(rings, N);
  k_gen_rows<N, DefaultGeom<N>::FWD_E><<<(K * N + CF::ROWS - 1) / CF::ROWS, CF::THREADS, 0, st>>>(desc, rings, tw);
  k_gen_cols<N, TC, EC><<<dim3((N / 2) / TC, K), TC * ColCfg<N, EC>::T, 0, st>>>(desc, tw);
  return hipGetLastError();
}

hipError_t launch_generate(int N, const CreatureDev* desc, RingParams* rings, int K, const float2* tw,
                           hipStream_t st) {
  switch (N) {
    case 16: return launch_generate_n<16>(desc, rings, K, tw, st);
    case 32: return launch_generate_n<32>(desc, rings, K, tw, st);
    case 64: return launch_generate_n<64>(desc, rings, K, tw, st);
    case 128: return launch_generate_n<128>(desc, rings, K, tw, st);
    case 256: return launch_generate_n<256>(desc, rings, K, tw, st);
    case 512: return launch_generate_n<512>(desc, rings, K, tw, st);
    case 1024: return launch_generate_n<1024>(desc, rings, K, tw, st);
    case 2048: return launch_generate_n<2048>(desc, rings, K, tw, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_seed(const CreatureDev* desc, int slot, size_t cells, uint32_t side, uint32_t origin,
                       uint64_t seed, uint64_t creature, hipStream_t st) {
  const size_t blocks = (cells + 255) / 256;
  k_gen_seed<<<(unsigned)(blocks < 65536 ? blocks : 65536), 256, 0, st>>>(desc, slot, side, origin, seed, creature);
  return hipGetLastError();
}

}  // namespace cn
