// Forward DiffusionLenia step kernels for gfx950.
//
// One step (reference phase sequence convolution -> growth -> diffusion -> copy,
// ml-rs/orchestrator/mod.rs:348-386) is four launches per grid-size class:
//   k_row_fwd         real rows -> half spectra (R2C via an N/2-point complex FFT)
//   k_col_fused       per source channel and column tile: forward column FFT, then for every
//                     kernel reading that source: product with the kernel spectrum and inverse
//                     column FFT (the spectrum stays in registers between kernels)
//   k_row_inv_growth  per row: C2R inverse row FFT of every kernel's result, growth bump and
//                     per-destination weighted sum, affinity exp (growth.rs, compute_1024.wgsl:299-325)
//   k_diffusion       halo-tiled 3x3 Z sum and mass-conserving gather (compute_1024.wgsl:341-413)
// The copy phase is a pointer swap (ping-pong state).
#include <hip/hip_runtime.h>
#include "cfg.h"
#include "fft.h"
#include "bufio.h"
#include "launch.h"
#include "mass.h"
#include "realfft.h"
#include "segio.h"
#include "types.h"

namespace cn {

// ---------------------------------------------------------------------------------------
// Row forward: state rows (real, N) -> half spectrum rows (N/2 complex, slot 0 = (X0, X_{N/2}))
// ---------------------------------------------------------------------------------------
template <int N, int ER>
__global__ void __launch_bounds__((RowCfg<N, ER>::THREADS))
k_row_fwd(const CreatureDev* __restrict__ descs, int sin, const float2* __restrict__ tw) {
  using CF = RowCfg<N, ER>;
  constexpr int L = CF::L, E = CF::E, T = CF::T, ROWS = CF::ROWS;
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N) return;
  const int nrows = d.C * N;
  const int row0 = blockIdx.x * ROWS;
  if (row0 >= nrows) return;
  const int lane = threadIdx.x / T, t = threadIdx.x % T;
  int row = row0 + lane;
  const bool valid = row < nrows;
  if (!valid) row = nrows - 1;

  gptr<const float2> in = (gptr<const float2>)(slot_ptr(d, sin) + (size_t)row * N);
  float2 v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = in[in_pos<L, E, false>(t, i)];

  // Twiddles in LDS (the row's own loads and stores do not wait behind them); one-wavefront
  // rows exchange at wavefront level, so the fill gets a block barrier of its own.
  using TF = TwLds<L, E, false>;
  __shared__ float2 twl[TF::size + SplitTw<N>::size];
  {
    TwFill<TF, SplitTw<N>, CF::THREADS> twf;
    twf.load(tw, threadIdx.x);
    twf.store(twl, threadIdx.x);
  }
  __syncthreads();
  __shared__ float2 lds[LdsMap<L, ROWS, false>::size];
  fft_run<L, E, ROWS, false, false, kTwN, 0, true>(v, lds, t, lane, twl);

  // Split the packed transform Z of z[n] = x[2n] + i x[2n+1] into the real row's spectrum.
  r2c_finish<N, E, ROWS, true>(v, lds, t, lane, twl + TF::size,
                               c64(d.spec) + (size_t)row * L, valid);
}

// ---------------------------------------------------------------------------------------
// Fused column pass: forward column FFT of one source tile, then per kernel of that source:
// product with the kernel's half spectrum and inverse column FFT.  SPLIT: one block per
// (tile, kernel) instead of per (tile, source), each redoing its source's forward transform —
// K instead of C blocks per tile, for grids too small to fill the GPU by sources.
//
// Memory: the source spectrum is read from its row-major plane and the results are written in
// the result layout (res_index: row-major, or column tiles whose 4 columns and 4 consecutive
// rows form one 128-byte line).  Every load and store goes through a buffer resource of its
// plane with a per-thread 32-bit offset and a compile-time per-element offset (bufio.h).
// ---------------------------------------------------------------------------------------
template <int N, int TC, int EC, bool SPLIT>
__global__ void __launch_bounds__((TC * ColCfg<N, EC>::T)) __attribute__((amdgpu_waves_per_eu(4)))
k_col_fused(const CreatureDev* __restrict__ descs, const float2* __restrict__ tw, int skip_seg) {
  constexpr int L = N, E = ColCfg<N, EC>::E, HALF = N / 2;
  constexpr uint32_t kPlaneBytes = (uint32_t)N * HALF * 8, kRealBytes = (uint32_t)(HALF + 1) * HALF * 4;
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N || (skip_seg && d.seg_g > 0)) return;  // skip_seg: the segmented pass covers those
  int c, kb, ke;
  if constexpr (SPLIT) {
    if ((int)blockIdx.y >= d.K) return;
    c = d.src_of[blockIdx.y];
    kb = 0;
    ke = 1;
  } else {
    c = blockIdx.y;
    if (c >= d.C) return;
    kb = d.src_off[c];
    ke = d.src_off[c + 1];
    if (kb == ke) return;
  }
  auto kernel_of = [&](int kk) { return SPLIT ? (int)blockIdx.y : d.src_k[kk]; };
  const int x0 = xcd_tile<(N / 2) / TC>(blockIdx.x) * TC;
  const int lane = threadIdx.x % TC;
  int t = threadIdx.x / TC;
  const int x = x0 + lane;

  // Element i of a thread sits at row t + (its position at t = 0): one VGPR offset per thread
  // (row t, column x) and one constant per element.
  const BufRes sp = make_res((const char*)d.spec + (size_t)c * kPlaneBytes, kPlaneBytes);
  float2 v[E];
  {
    const uint32_t vo = (uint32_t)(t * HALF + x) * 8u;
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = buf_load_c64(sp, vo, (uint32_t)in_pos<L, E, false>(0, i) * HALF * 8u);
  }

  // Kernel spectra are prefetched one kernel ahead: the loads of kernel k+1's tile are in
  // flight while kernel k's inverse transform runs (the first ones during the forward one).
  // Kernels with a real spectrum that is even in ky are read from the compact real copy
  // (a quarter of the bytes) and enter a real product (2 multiplies per value instead of 6).
  // Rows y > N/2 of a compact spectrum read row N - y; every element's side of N/2 is fixed
  // at compile time (element offsets are multiples of N/E, as is N/2).
  auto load_kernel = [&](int k, float2* out) {
    if (d.ksym[k]) {
      const BufRes kr = make_res(d.kreal + (size_t)k * (HALF + 1) * HALF, kRealBytes);
      const uint32_t vb = (uint32_t)((x >> 2) * (HALF + 1) * 4 + (x & 3)) * 4u;
      const uint32_t vp = vb + (uint32_t)t * 16u, vn = vb + (uint32_t)(N - t) * 16u;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int ci = out_pos<L, E, false>(0, i);
        out[i].x = ci < HALF ? buf_load_f32(kr, vp, (uint32_t)ci * 16u) : buf_load_f32(kr, vn - (uint32_t)ci * 16u, 0);
      }
    } else {
      const BufRes kh = make_res((const char*)d.khalf + (size_t)k * kPlaneBytes, kPlaneBytes);
      const uint32_t vo = (uint32_t)(t * HALF + x) * 8u;
#pragma unroll
      for (int i = 0; i < E; ++i) out[i] = buf_load_c64(kh, vo, (uint32_t)out_pos<L, E, false>(0, i) * HALF * 8u);
    }
  };
  auto product = [&](int k, const float2* nkv, float2* wv) {
    if (d.ksym[k]) {
#pragma unroll
      for (int i = 0; i < E; ++i) wv[i] = make_float2(v[i].x * nkv[i].x, v[i].y * nkv[i].x);
    } else {
#pragma unroll
      for (int i = 0; i < E; ++i) wv[i] = c_mul(v[i], nkv[i]);
    }
  };
  // Twiddles of both plans in LDS (see TwLds): their loads go out right behind the spectrum's
  // and ahead of the first kernel's, the LDS stores wait for them alone, and the first
  // exchange barrier orders the stores.
  using TF = TwLds<L, E, false>;
  using TI = TwLds<L, E, true>;
  __shared__ float2 twl[TF::size + TI::size + 1];
  TwFill<TF, TI, TC * ColCfg<N, EC>::T> twf;
  twf.load(tw, threadIdx.x);
  float2 nk[E];
  load_kernel(kernel_of(kb), nk);
  twf.store(twl, threadIdx.x);
  __shared__ float2 lds[LdsMap<L, TC, true>::size];
  fft_run<L, E, TC, true, false, kTwN, 0, true>(v, lds, t, lane, twl);

  // Tile 0, column 0 carries two real columns packed as (DC column) + i (Nyquist column).
  // Per kernel, lane 0 splits its spectrum F into A = FFT(DC column) and B = FFT(Nyquist
  // column) through the exchange buffer (free between transforms), multiplies each by its own
  // kernel column and re-packs A K_dc + i B K_nyq.
  //
  // The first product is formed before the loop: the entry path then has no load of nk in
  // flight at the loop head, and the wait for a later kernel's prefetch (issued before the
  // previous kernel's stores) leaves those stores in flight.  Formed only inside the loop, the
  // head's wait would also cover the entry path, where the prefetch is the newest memory
  // operation, and would wait for everything.
  const bool edge = (x0 == 0);
  float2 w[E];
  product(kernel_of(kb), nk, w);
  // Result plane offsets (res_index): row t of column x, and each element's row offset.
  const uint32_t vct = kColTiled ? (uint32_t)(((x >> 2) * N + t) * 4 + (x & 3)) * 8u : (uint32_t)(t * HALF + x) * 8u;
  constexpr uint32_t kRowStride = kColTiled ? 32u : (uint32_t)HALF * 8u;
  for (int kk = kb; kk < ke; ++kk) {
    const int k = kernel_of(kk);
    t = launder(t);
    if (kk > kb) product(k, nk, w);
    if (kk + 1 < ke) load_kernel(kernel_of(kk + 1), nk);
    if (edge) {
      // Lane 0 splits the packed DC/Nyquist column; its kernel factors (khalf column 0, knyq)
      // are loaded a few at a time so that the loads overlap.
      __syncthreads();
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < E; ++i) lds[LdsMap<L, TC, true>::at(out_pos<L, E, false>(t, i), 0)] = v[i];
      }
      __syncthreads();
      if (lane == 0) {
        gptr<const float2> kh = c64(d.khalf) + (size_t)k * N * HALF;
        gptr<const float2> kn = c64(d.knyq) + (size_t)k * N;
        constexpr int G4 = 2;
#pragma unroll
        for (int i0 = 0; i0 < E; i0 += G4) {
          float2 kd[G4], kq[G4];
#pragma unroll
          for (int u = 0; u < G4; ++u) {
            const int y = out_pos<L, E, false>(t, i0 + u);
            kd[u] = kh[(size_t)y * HALF];
            kq[u] = kn[y];
          }
#pragma unroll
          for (int u = 0; u < G4; ++u) {
            const int i = i0 + u;
            const int y = out_pos<L, E, false>(t, i);
            const float2 F = v[i];
            const float2 G = lds[LdsMap<L, TC, true>::at((N - y) & (N - 1), 0)];
            const float2 a = make_float2(0.5f * (F.x + G.x), 0.5f * (F.y - G.y));  // (F + conj G)/2
            const float2 dd = make_float2(F.x - G.x, F.y + G.y);                    // F - conj G
            const float2 b = make_float2(0.5f * dd.y, -0.5f * dd.x);                // dd / (2i)
            const float2 pa = c_mul(a, kd[u]);
            const float2 pb = c_mul(b, kq[u]);
            w[i] = make_float2(pa.x - pb.y, pa.y + pb.x);  // pa + i*pb
          }
        }
      }
    }
    // The pass is bound by its store stream: from the inverse transform on, the wave runs at
    // raised priority, so that waves with stores to issue are not held back by the others'
    // loads and forward work (raising it around the spectrum loads instead made the pass
    // slower).
    __builtin_amdgcn_s_setprio(1);
    fft_run<L, E, TC, true, true, kTwN, 0, true>(w, lds, t, lane, twl + TF::size);
    const BufRes uo = make_res((const char*)d.uhat + (size_t)k * kPlaneBytes, kPlaneBytes);
#pragma unroll
    for (int i = 0; i < E; ++i) buf_store_c64(w[i], uo, vct, (uint32_t)out_pos<L, E, true>(0, i) * kRowStride);
    __builtin_amdgcn_s_setprio(0);
  }
}

// ---------------------------------------------------------------------------------------
// Segmented column pass (overlap-save along y), for creatures whose kernels all have real, even
// spectra with column profiles of support |y| <= R (segment.hip): the convolution of each
// spectrum column with its profile is computed on G segments of M rows.  Segment s produces the
// output rows [o_s, o_{s+1}), o_s = s N / G, from the M input rows starting at o_s - R (rows
// wrap around the torus): forward M-point transform, product with the kernel's segment table
// H_k(m, kx) (real, even in m; the 1/M of the inverse folded in), inverse transform, and of its
// M outputs the rows R .. R + o_{s+1} - o_s - 1 are stored (the others hold wrapped terms).
// A block takes seg_tc(N) columns (16 or 32) of one segment of one source: its loads and stores
// are whole 128-byte lines (16 consecutive complex values per row), and its job is short, so
// that the blocks resident on a CU are at different phases of their jobs.
//
// Column kx = 0 packs the DC and Nyquist columns (X_0, X_{N/2}) as z = a + i b: with Z its
// transform, A H0 + i B HN = Z (H0 + HN)/2 + conj(Z(-m)) (H0 - HN)/2 (H0, HN real), the mirror
// Z(-m) taken through the exchange buffer; HN is the table's column N/2.
// ---------------------------------------------------------------------------------------
template <int N, int M, int E>
__global__ void __launch_bounds__((seg_tc(N) * (M / E))) __attribute__((amdgpu_waves_per_eu(4)))
k_col_segments(const CreatureDev* __restrict__ descs, const float2* __restrict__ tw) {
  constexpr int TC = seg_tc(N), HALF = N / 2, W = HALF + 1, NT = HALF / TC, MH = M / 2;
  constexpr uint32_t kPlaneBytes = (uint32_t)N * HALF * 8, kTabBytes = (uint32_t)(MH + 1) * W * 4;
  using LM = LdsMap<M, TC, true>;
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N) return;
  const int c = blockIdx.y;
  if (c >= d.C) return;
  const int kb = d.src_off[c], ke = d.src_off[c + 1];
  if (kb == ke) return;
  const int G = d.seg_g, R = d.seg_r;
  // Jobs (tile, segment), segments fastest, dealt so that each XCD takes a contiguous run: the
  // rows two neighbouring segments share are read through the same L2.
  const int nb = (int)gridDim.x, bx = (int)blockIdx.x, gmax = nb / NT;
  const int job = (nb % 8 == 0) ? (bx % 8) * (nb / 8) + bx / 8 : bx;
  const int tile = job / gmax, seg = job % gmax;
  if (seg >= G) return;
  const int lane = (int)threadIdx.x % TC, x = tile * TC + lane;
  int t = (int)threadIdx.x / TC;
  const int o0 = seg * N / G, rows = (seg + 1) * N / G - o0, base = o0 - R;

  using IO = SegIO<N, M, E>;
  float2 v[E];
  IO::load((const char*)d.spec + (size_t)c * kPlaneBytes, base, t, x, v);
  // Table factor of slot i (frequency m = out_pos(t, i), row min(m, M - m)).
  auto load_h = [&](int k, float* out) { IO::table(d.kseg + (size_t)k * (MH + 1) * W, t, x, out); };
  using TF = TwLds<M, E, false>;
  using TI = TwLds<M, E, true>;
  __shared__ float2 twl[TF::size + TI::size + 1];
  TwFill<TF, TI, TC * (M / E)> twf;
  twf.load(tw, threadIdx.x);
  float hk[E];
  load_h(d.src_k[kb], hk);
  twf.store(twl, threadIdx.x);
  __shared__ float2 lds[LM::size];
  fft_run<M, E, TC, true, false, kTwN, 0, true>(v, lds, t, lane, twl);

  const bool edge = tile == 0;
  for (int kk = kb; kk < ke; ++kk) {
    const int k = d.src_k[kk];
    t = launder(t);
    float2 w[E];
#pragma unroll
    for (int i = 0; i < E; ++i) w[i] = make_float2(v[i].x * hk[i], v[i].y * hk[i]);
    if (edge) {
      __syncthreads();
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < E; ++i) lds[LM::at(out_pos<M, E, false>(t, i), 0)] = v[i];
      }
      __syncthreads();
      if (lane == 0) {
        const BufRes kr = make_res(d.kseg + (size_t)k * (MH + 1) * W, kTabBytes);
#pragma unroll
        for (int i = 0; i < E; ++i) {
          const int m = out_pos<M, E, false>(t, i), mm = m <= MH ? m : M - m;
          const float hn = buf_load_f32(kr, (uint32_t)(mm * W + HALF) * 4u, 0);
          const float p = 0.5f * (hk[i] + hn), q = 0.5f * (hk[i] - hn);
          const float2 F = v[i], Gm = lds[LM::at((M - m) & (M - 1), 0)];
          w[i] = make_float2(F.x * p + Gm.x * q, F.y * p - Gm.y * q);
        }
      }
    }
    if (kk + 1 < ke) load_h(d.src_k[kk + 1], hk);
    __builtin_amdgcn_s_setprio(1);
    fft_run<M, E, TC, true, true, kTwN, 0, true>(w, lds, t, lane, twl + TF::size);
    IO::store((char*)d.uhat + (size_t)k * kPlaneBytes, w, t, x, o0, R, rows);
    __builtin_amdgcn_s_setprio(0);
  }
}

// ---------------------------------------------------------------------------------------
// Row inverse (C2R) of every kernel's result + growth + weighted sum + affinity exp.
// ---------------------------------------------------------------------------------------
template <int N, int ER, bool WRITE_U>
__global__ void __launch_bounds__((RowCfg<N, ER>::THREADS)) __attribute__((amdgpu_waves_per_eu(4)))
k_row_inv_growth(const CreatureDev* __restrict__ descs, const float2* __restrict__ tw, int islot) {
  using CF = RowCfg<N, ER>;
  constexpr int L = CF::L, E = CF::E, T = CF::T, ROWS = CF::ROWS;
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N) return;
  const int c = blockIdx.y;  // destination channel
  if (c >= d.C) return;
  const int lane = threadIdx.x / T, t0 = threadIdx.x % T;
  const int y = blockIdx.x * ROWS + lane;  // ROWS divides N
  __shared__ float2 lds[LdsMap<L, ROWS, false>::size];
  using TI = TwLds<L, E, true>;
  __shared__ float2 twl[TI::size + SplitTw<N>::size];  // LDS twiddles, as in k_row_fwd
  {
    TwFill<TI, SplitTw<N>, CF::THREADS> twf;
    twf.load(tw, threadIdx.x);
    twf.store(twl, threadIdx.x);
  }
  __syncthreads();
  const float temp = d.temp;
  {
    float2 acc[E];  // sum_k 2 w_k exp(...); the affinity is acc - wsum
#pragma unroll
    for (int i = 0; i < E; ++i) acc[i] = make_float2(0.f, 0.f);
    float wsum = 0.f;
    const int kb = d.dst_off[c], ke = d.dst_off[c + 1];
    // Convolution results in the column pass's result layout (res_index).
    auto uhat_plane = [&](int k) { return (const void*)((const float2*)d.uhat + (size_t)k * N * L); };
    // Each kernel's row is prefetched one kernel ahead (loads in flight during the transform).
    // Every lane loads only its own elements and takes the partners through LDS.  With 16
    // values per thread the prefetch registers would spill; the row is loaded at the start of
    // its iteration instead (a one-wavefront row has no block barrier to wait on).
    constexpr bool kWave = wave_rows<L, E>();
    constexpr bool kPrefetch = !(kWave && E >= 16);
    float2 xp[E], xq[E];
    if (kPrefetch && kb < ke) c2r_load_own_ct<N, E>(uhat_plane(d.dst_k[kb]), y, t0, xp);
    for (int kk = kb; kk < ke; ++kk) {
      const int k = d.dst_k[kk];
      const int t = launder(t0);
      // G = 2 exp(-(u-mu)^2 / (2 sigma^2)) - 1 enters a weighted sum, so its absolute error is
      // what matters: 2^(d (d c)) with c = -log2(e) / (2 sigma^2) is within ~1e-7 of the exact
      // value (the argument's relative error ~3 * 2^-24 |p| is damped by 2^-|p|).  The "- 1"
      // terms are folded into one subtraction of sum_k w_k per destination.
      const float mu = d.mu[k], wk = d.w[k];
      const float cexp = d.inv_s2[k] * -0.72134752044448170368f;  // -log2(e) / 2 / sigma^2
      const float w2 = 2.0f * wk;
      wsum += wk;
      float2 v[E];
      // Row loads at raised priority: without a prefetch they gate this wave's progress, and
      // issued ahead of the other waves' transform work they overlap it (2 % on the pass).
      __builtin_amdgcn_s_setprio(1);
      if constexpr (!kPrefetch) c2r_load_own_ct<N, E>(uhat_plane(k), y, t, xp);
      __builtin_amdgcn_s_setprio(0);
      c2r_partner<N, E, ROWS>(xp, lds, t, lane, xq);
      c2r_form<N, E, true>(xp, xq, t, twl + TI::size, v, mu);  // the transform yields u - mu
      __builtin_amdgcn_s_setprio(1);
      if (kPrefetch && kk + 1 < ke) c2r_load_own_ct<N, E>(uhat_plane(d.dst_k[kk + 1]), y, t, xp);
      __builtin_amdgcn_s_setprio(0);
      fft_run<L, E, ROWS, false, true, kTwN, 0, true>(v, lds, t, lane, twl);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int n = out_pos<L, E, true>(t, i);
        const float d0 = v[i].x, d1 = v[i].y;  // u[2n] - mu, u[2n+1] - mu
        if constexpr (WRITE_U) {
          ((gptr<float2>)to_global(d.u + (((size_t)islot * d.K + k) * N + y) * N))[n] = make_float2(d0 + mu, d1 + mu);
        }
        acc[i].x = fmaf(w2, __builtin_amdgcn_exp2f(d0 * (d0 * cexp)), acc[i].x);
        acc[i].y = fmaf(w2, __builtin_amdgcn_exp2f(d1 * (d1 * cexp)), acc[i].y);
      }
    }
    gptr<float2> Ao = (gptr<float2>)to_global(d.A + (((size_t)islot * d.C + c) * N + y) * N);
    const int t = launder(t0);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int n = out_pos<L, E, true>(t, i);
      Ao[n] = make_float2(exp_fast(temp * (acc[i].x - wsum)), exp_fast(temp * (acc[i].y - wsum)));
    }
  }
}

// ---------------------------------------------------------------------------------------
// Row inverse of every kernel's result for two rows at once + growth + weighted sum + exp.
//
// Rows y0 and y0 + 1 of a kernel's result are real, so their half spectra Ha, Hb (N/2 complex,
// slot 0 packing the real bins X[0], X[N/2]) extend to Hermitian spectra Ua, Ub and one
// N-point complex inverse transform of Y = Ua + i Ub gives y = ua + i ub.  Against two C2R
// rows (an N/2-point transform and the real split each) this drops the split's twiddle
// products and pre-pass at the price of one radix-2 stage's worth of transform work.
//
// Thread t of a row pair holds Y at kx = t + m T2 (m = 0..E-1, T2 = N/E), i.e. slots m < E/2 at
// kx < N/2 and slots m >= E/2 at mirrors N - kx = (T2 - t) + (E-1-m) T2 < N/2 of thread T2 - t's
// first half.  Each thread loads its first-half elements of both rows once and takes the
// mirrors from LDS:
//   kx < N/2:  Y = Ha[kx] + i Hb[kx]             = (a.x - b.y, a.y + b.x)
//   kx > N/2:  Y = conj(Ha[N-kx]) + i conj(Hb[N-kx]) = (a.x + b.y, b.x - a.y)
//   kx = 0, N/2 (thread 0): Y = (Ha[0].x, Hb[0].x), (Ha[0].y, Hb[0].y).
// ---------------------------------------------------------------------------------------
// Where the pair form runs (measured at the benchmark shapes, same build, same run): 2048^2
// C3K9 inverse-row pass 47.3 us (pairs, 8 values per thread) against 49.4 us (C2R rows, 16
// values per thread, the faster C2R geometry); at 1024^2 C3K9 the C2R rows are faster (rollout
// 11.06 against 11.28 ms) and at 512^2 C5K25 the two are on par, so smaller grids keep the C2R
// rows (a one-wavefront row needs no block barrier; a 2048-point pair at 16 values per thread
// spans two wavefronts and needs six per kernel).
template <int N>
constexpr bool inv_row_pairs() {
  return N >= 2048;
}

template <int N, int ER>
struct PairCfg {
  static constexpr int E = ER < N ? ER : N;
  static constexpr int T = N / E;  // threads per row pair
  static constexpr int PAIRS = T >= 128 ? 1 : ((256 / T) < N / 2 ? (256 / T) : N / 2);
  static constexpr int THREADS = PAIRS * T;
  static constexpr int BLOCKS = N / (2 * PAIRS);  // blocks per plane
  static_assert(N % (2 * PAIRS) == 0, "row pairs per block must divide the plane");
};

template <int N, int ER, bool WRITE_U>
__global__ void __launch_bounds__((PairCfg<N, ER>::THREADS)) __attribute__((amdgpu_waves_per_eu(4)))
k_row_inv_growth_pair(const CreatureDev* __restrict__ descs, const float2* __restrict__ tw, int islot) {
  using CF = PairCfg<N, ER>;
  constexpr int L = N / 2, E = CF::E, T = CF::T, PAIRS = CF::PAIRS, H = E / 2;
  static_assert(E % 2 == 0 && (L / T) == H, "a thread's first-half slots must be E/2");
  using LM = LdsMap<N, PAIRS, false>;
  constexpr bool kWave = wave_rows<N, E>();
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N) return;
  const int c = blockIdx.y;  // destination channel
  if (c >= d.C) return;
  const int pr = threadIdx.x / T, t0 = threadIdx.x % T;
  const int y0 = (blockIdx.x * PAIRS + pr) * 2;
  __shared__ float2 lds[LM::size];
  using TI = TwLds<N, E, true>;
  __shared__ float2 twl[TI::size + 1];
  {
    TwFill<TI, TwNone, CF::THREADS> twf;
    twf.load(tw, threadIdx.x);
    twf.store(twl, threadIdx.x);
  }
  __syncthreads();
  // The last pass's two-factor twiddles are the same for every kernel of the destination: form
  // them once (the pass kernel's occupancy is bound by its LDS, which leaves these registers).
  TwRegs<N, E, true> twr;
  twr.init(twl, launder(t0));
  const float temp = d.temp;
  float2 acc[E];  // (row y0, row y0 + 1) of sum_k 2 w_k exp(...) at kx-slot outputs
#pragma unroll
  for (int i = 0; i < E; ++i) acc[i] = make_float2(0.f, 0.f);
  float wsum = 0.f;
  const int kb = d.dst_off[c], ke = d.dst_off[c + 1];
  // First-half elements kx = t + m T (m < E/2) of rows y0, y0 + 1 of kernel k's result plane.
  auto load_own = [&](int k, int t, float2* xa, float2* xb) {
    const void* plane = (const void*)((const float2*)d.uhat + (size_t)k * N * L);
    if constexpr (!kColTiled) {
      const BufRes r = make_res(plane, (uint32_t)N * L * 8u);
      const uint32_t vo = (uint32_t)(y0 * L + t) * 8u;
#pragma unroll
      for (int m = 0; m < H; ++m) {
        xa[m] = buf_load_c64(r, vo, (uint32_t)(m * T) * 8u);
        xb[m] = buf_load_c64(r, vo, (uint32_t)(L + m * T) * 8u);
      }
    } else {
      gptr<const float2> X = c64(plane);
#pragma unroll
      for (int m = 0; m < H; ++m) {
        xa[m] = X[res_index(N, y0, t + m * T)];
        xb[m] = X[res_index(N, y0 + 1, t + m * T)];
      }
    }
  };
  // At 8 values per thread (without the u stores) the kernel-loop-invariant index arithmetic
  // fits the registers the LDS-bound occupancy leaves (71 VGPRs, 10 % fewer loop instructions);
  // otherwise the lane index is laundered per kernel so that it is recomputed instead of spilled.
  constexpr bool kHoist = E <= 8 && !WRITE_U;
  const int tl = launder(t0);
  for (int kk = kb; kk < ke; ++kk) {
    const int k = d.dst_k[kk];
    const int t = kHoist ? tl : launder(t0);
    const float mu = d.mu[k], wk = d.w[k];
    const float cexp = d.inv_s2[k] * -0.72134752044448170368f;  // -log2(e) / 2 / sigma^2
    const float w2 = 2.0f * wk;
    wsum += wk;
    float2 xa[H], xb[H];
    __builtin_amdgcn_s_setprio(1);
    load_own(k, t, xa, xb);
    __builtin_amdgcn_s_setprio(0);
    // Mirror exchange: row r's half spectrum at padded position r L + kx of the pair's buffer.
    xsync<kWave>();
    if constexpr (T % 16 == 0) {
      // positions t + c with c a multiple of 16: one base, immediate offsets
      float2* p = lds + LM::at(t, pr);
#pragma unroll
      for (int m = 0; m < H; ++m) {
        p[LM::off(m * T)] = xa[m];
        p[LM::off(L + m * T)] = xb[m];
      }
    } else {
#pragma unroll
      for (int m = 0; m < H; ++m) {
        lds[LM::at(t + m * T, pr)] = xa[m];
        lds[LM::at(L + t + m * T, pr)] = xb[m];
      }
    }
    xsync<kWave>();
    float2 v[E];
    static_for<0, E>([&](auto I) {
      constexpr int i = decltype(I)::value, ci = in_off<N, E, true>(i), m = ci / T;
      if constexpr (m < H) {
        const float2 a = xa[m], b = xb[m];
        const float2 g = make_float2(a.x - b.y, a.y + b.x);
        if constexpr (m == 0) {
          // The rows' DC bins less mu: the transform then gives u - mu (a constant row
          // spectrum at kx = 0 adds that constant to every output), one subtraction per row
          // pair instead of one per output.
          v[i] = (t == 0) ? make_float2(a.x - mu, b.x - mu) : g;
        } else {
          v[i] = g;
        }
      } else {
        // N - kx = L - t - (m - H) T; thread 0's slot m = H is the packed Nyquist bin.
        const int q = (L - t - (m - H) * T) & (L - 1);
        const float2 a = lds[LM::at(q, pr)], b = lds[LM::at(L + q, pr)];
        const float2 g = make_float2(a.x + b.y, b.x - a.y);
        if constexpr (m == H) {
          v[i] = (t == 0) ? make_float2(xa[0].y, xb[0].y) : g;
        } else {
          v[i] = g;
        }
      }
    });
    fft_run<N, E, PAIRS, false, true, kTwN, 0, true, true>(v, lds, t, pr, twl, twr.w);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int x = out_pos<N, E, true>(t, i);
      if constexpr (WRITE_U) {
        gptr<float> uo = to_global(d.u + (((size_t)islot * d.K + k) * N + y0) * N);
        uo[x] = v[i].x + mu;
        uo[N + x] = v[i].y + mu;
      }
      const float d0 = v[i].x, d1 = v[i].y;  // u - mu
      acc[i].x = fmaf(w2, __builtin_amdgcn_exp2f(d0 * (d0 * cexp)), acc[i].x);
      acc[i].y = fmaf(w2, __builtin_amdgcn_exp2f(d1 * (d1 * cexp)), acc[i].y);
    }
  }
  gptr<float> Ao = to_global(d.A + (((size_t)islot * d.C + c) * N + y0) * N);
  const int t = launder(t0);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int x = out_pos<N, E, true>(t, i);
    Ao[x] = exp_fast(temp * (acc[i].x - wsum));
    Ao[N + x] = exp_fast(temp * (acc[i].y - wsum));
  }
}

// ---------------------------------------------------------------------------------------
// Diffusion: Z = box3(A) (periodic), new = A * box3(state / Z).  Halo-2 tile in LDS.
//
// Mass conservation.  Exactly, sum_p new_p = sum_n (s_n / Z_n) * sum_{p in N(n)} A_p
// = sum_n s_n.  Plain FP32 breaks the identity by a systematic ~5e-8 per step (the reference's
// f32 CPU step drifts the same way): Z_n is not exactly the sum of the A_p that later receive
// q_n, and in smooth regions these roundings repeat with the same sign.  Here Z and the gather
// are FP64 sums of the f32 tile values (rounding 2^-53; FP64 adds issue at the FP32 rate, and
// the LDS tile stays f32), q_n = s_n / Z_n is a fast f32 quotient, and each site keeps the FP64 residual
// r_n = s_n - q_n Z_n of its division and adds it to its own output, so whatever q_n is,
// sum_p new_p = sum_n (q_n Z_n + r_n) = sum_n s_n up to FP64 rounding.  What is left is the final
// rounding of each output to f32 (round to nearest, no systematic sign).
//
// Layout: one wavefront per tile of 60 columns x TH rows, no LDS and no barriers.  Lane l holds
// column x0 - 2 + l (64 columns: the tile plus a 2-column halo on each side).  Horizontal 3-sums
// take the neighbours' values with DPP wave shifts; vertical 3-sums are rolling registers while
// the wave walks down the rows (A rows y0-2 .. y0+TH+1, state rows y0-1 .. y0+TH).  Z and q are
// valid on lanes 1..62, the gather on lanes 2..61, which store.  Tiles are dealt to blocks so
// that each XCD (its own L2) gets a contiguous band of tiles: neighbouring tiles share halo
// lines through the same L2.
// ---------------------------------------------------------------------------------------
constexpr int kDiffTW = 60, kDiffThreads = 256;  // output columns per wave, 4 waves per block

// IEEE round-to-nearest f32 square root of x in [0, 1].  The hardware square root is within
// 1 ulp; the correctly rounded result t is the candidate in {s - ulp, s, s + ulp} whose rounding
// interval holds sqrt(x), i.e. lo^2 <= x <= hi^2 for the midpoints lo, hi to its neighbours.
// Midpoints have 25 significant bits, so their squares and the comparisons are exact in FP64.
__device__ __forceinline__ float sqrt_rn01(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  if (!(x > 0.0f)) return 0.0f;
  const float dn = __uint_as_float(__float_as_uint(s) - 1u), up = __uint_as_float(__float_as_uint(s) + 1u);
  const double xd = (double)x;
  const double lo = 0.5 * ((double)dn + (double)s), hi = 0.5 * ((double)s + (double)up);
  if (xd < lo * lo) return dn;
  if (xd > hi * hi) return up;
  return s;
}

// One colour byte: u32(sqrt(clamp(1.5 v, 0, 1)) * 255) with IEEE f32 sqrt and truncation
// (compute_1024.wgsl:438-440).  NaN maps to 0.
__device__ __forceinline__ unsigned char render_byte(float v) {
  const float x = fminf(fmaxf(v * 1.5f, 0.0f), 1.0f);
  return (unsigned char)(unsigned int)(sqrt_rn01(x) * 255.0f);
}

// x(l-1) + x(l) + x(l+1) in FP64 across the wavefront (lanes 0 and 63 see a 0 neighbour).
__device__ __forceinline__ double hsum3(float x) {
  const float lo = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, false));
  const float hi = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, false));
  return (double)lo + (double)x + (double)hi;
}

template <int TH>
__global__ void __launch_bounds__(kDiffThreads)
k_diffusion(const CreatureDev* __restrict__ descs, int sin, int sout, int N, int render, int islot) {
  constexpr int OW = kDiffTW;
  const CreatureDev& d = descs[blockIdx.z];
  if (d.N != N) return;
  const int c = blockIdx.y;
  if (c >= d.C) return;
  const int tiles_x = (N + OW - 1) / OW, tiles_y = (N + TH - 1) / TH;
  const int nblk = gridDim.x, bx = (int)blockIdx.x;
  const int blk = (nblk % 8 == 0) ? (bx % 8) * (nblk / 8) + bx / 8 : bx;
  const int tile = blk * (kDiffThreads / 64) + (int)(threadIdx.x >> 6);
  if (tile >= tiles_x * tiles_y) return;  // a whole wavefront; the kernel has no barriers
  const int l = threadIdx.x & 63;
  const int y0 = (tile / tiles_x) * TH, x0 = (tile % tiles_x) * OW;
  const int mask = N - 1;
  const int xl = x0 - 2 + l;             // this lane's column (unwrapped)
  const int gx = xl & mask;
  const bool store = l >= 2 && l < 2 + OW && xl < N;
  const size_t plane = (size_t)N * N;
  gptr<const float> A = to_global((const float*)d.A) + ((size_t)islot * d.C + c) * plane + gx;
  gptr<const float> s = slot_ptr(d, sin) + c * plane + gx;
  gptr<float> o = slot_ptr(d, sout) + c * plane + xl;
  // Render epilogue (render.rs:145-153, compute_1024.wgsl:429-443): channel c < 3 writes its
  // byte of the packed 0x00RRGGBB pixel (little endian: R is byte 2, B byte 0).
  gptr<unsigned char> rgb = (render && c < 3 && d.render) ? to_global(reinterpret_cast<unsigned char*>(d.render)) : nullptr;

  // All rows of the tile are loaded up front (independent loads in flight together), then
  // step j computes q on row y0 - 1 + j; from j = 2 on it also finishes output row y0 + j - 2.
  float ar[TH + 4], sr[TH + 2];
#pragma unroll
  for (int j = 0; j < TH + 4; ++j) ar[j] = A[(size_t)((y0 - 2 + j) & mask) * N];
#pragma unroll
  for (int j = 0; j < TH + 2; ++j) sr[j] = s[(size_t)((y0 - 1 + j) & mask) * N];
  double h0 = hsum3(ar[0]);  // horizontal A sums of rows yq-1, yq
  double h1 = hsum3(ar[1]);
  double g0 = 0.0, g1 = 0.0;  // horizontal q sums of the two previous q rows
  float a0 = 0.f, a1 = 0.f;   // A of the two previous A rows loaded in the steps
  float r1 = 0.f;             // residual of the previous q row
#pragma unroll
  for (int j = 0; j < TH + 2; ++j) {
    const int yq = y0 - 1 + j;
    const float a2 = ar[j + 2];
    const float sv = sr[j];
    const double h2 = hsum3(a2);
    const double z = h0 + h1 + h2;
    const float q = sv * __builtin_amdgcn_rcpf((float)z);
    const float r = (float)fma(-(double)q, z, (double)sv);
    const double g2 = hsum3(q);
    if (j >= 2) {
      const int yo = yq - 1;
      const float v = (float)fma((double)a0, g0 + g1 + g2, (double)r1);
      if (store && yo < N) {
        o[(size_t)yo * N] = v;
        if (rgb) rgb[((size_t)yo * N + xl) * 4 + (2 - c)] = render_byte(v);
      }
    }
    h0 = h1;
    h1 = h2;
    g0 = g1;
    g1 = g2;
    a0 = a1;
    a1 = a2;
    r1 = r;
  }
}

// ---------------------------------------------------------------------------------------
// FP64 Neumaier-compensated mass per channel (mass.h).
// partial[(creature * 16 + c) * nblk + blockIdx.x] = (sum, compensation)
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kMassThreads)
k_mass(const CreatureDev* __restrict__ descs, int slot, double2* __restrict__ partial, int nblk) {
  const CreatureDev& d = descs[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= d.C) return;
  const size_t plane = (size_t)d.N * d.N;
  mass_block(d.state + ((size_t)slot * d.C + c) * plane, plane, (int)blockIdx.x, nblk,
             partial + ((size_t)blockIdx.z * kMaxChannels + c) * nblk + blockIdx.x);
}

// ---------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------
static hipError_t record(hipEvent_t ev, hipStream_t st) { return ev ? hipEventRecord(ev, st) : hipSuccess; }

template <int N, int TC, int EC>
static hipError_t launch_col(const StepLaunch& sl, hipStream_t st) {
  if constexpr (!col_tc_ok<N, TC, EC>()) {
    return hipErrorInvalidValue;
  } else {
    const int skip = sl.seg_g > 0 ? 1 : 0;
    if (sl.col_split) {
      const dim3 g_col((N / 2) / TC, sl.max_k, sl.count);
      k_col_fused<N, TC, EC, true><<<g_col, TC * ColCfg<N, EC>::T, 0, st>>>(sl.descs, sl.tw, skip);
    } else {
      const dim3 g_col((N / 2) / TC, sl.max_c, sl.count);
      k_col_fused<N, TC, EC, false><<<g_col, TC * ColCfg<N, EC>::T, 0, st>>>(sl.descs, sl.tw, skip);
    }
    return hipSuccess;
  }
}

template <int N, int EC>
static hipError_t launch_col_tc(const StepLaunch& sl, int tc, hipStream_t st) {
  switch (tc) {
    case 2: return launch_col<N, 2, EC>(sl, st);
    case 4: return launch_col<N, 4, EC>(sl, st);
    case 8: return launch_col<N, 8, EC>(sl, st);
    case 16: return launch_col<N, 16, EC>(sl, st);
    default: return hipErrorInvalidValue;
  }
}

template <int N, int ER>
static hipError_t launch_rows(const StepLaunch& sl, hipStream_t st, bool fwd) {
  using CF = RowCfg<N, ER>;
  if (fwd) {
    const dim3 g_row((sl.max_c * N + CF::ROWS - 1) / CF::ROWS, 1, sl.count);
    k_row_fwd<N, ER><<<g_row, CF::THREADS, 0, st>>>(sl.descs, sl.sin, sl.tw);
  } else {
    if constexpr (inv_row_pairs<N>()) {
      using PC = PairCfg<N, ER>;
      const dim3 g_inv(PC::BLOCKS, sl.max_c, sl.count);
      if (sl.write_u)
        k_row_inv_growth_pair<N, ER, true><<<g_inv, PC::THREADS, 0, st>>>(sl.descs, sl.tw, sl.islot);
      else
        k_row_inv_growth_pair<N, ER, false><<<g_inv, PC::THREADS, 0, st>>>(sl.descs, sl.tw, sl.islot);
    } else {
      const dim3 g_inv(CF::BLOCKS, sl.max_c, sl.count);
      if (sl.write_u)
        k_row_inv_growth<N, ER, true><<<g_inv, CF::THREADS, 0, st>>>(sl.descs, sl.tw, sl.islot);
      else
        k_row_inv_growth<N, ER, false><<<g_inv, CF::THREADS, 0, st>>>(sl.descs, sl.tw, sl.islot);
    }
  }
  return hipSuccess;
}

template <int N>
static hipError_t launch_rows_e(const StepLaunch& sl, hipStream_t st, bool fwd) {
  const int e = fwd ? (sl.fwd_e ? sl.fwd_e : DefaultGeom<N>::FWD_E) : (sl.inv_e ? sl.inv_e : DefaultGeom<N>::INV_E);
  switch (e) {
    case 8: return launch_rows<N, 8>(sl, st, fwd);
    case 16: return launch_rows<N, 16>(sl, st, fwd);
    default: return hipErrorInvalidValue;
  }
}

template <int N>
static hipError_t launch_col_segments(const StepLaunch& sl, hipStream_t st) {
  constexpr int M = seg_m(N), E = seg_e(N);
  if constexpr (M == 0) {
    return hipErrorInvalidValue;
  } else {
    const dim3 g((N / 2 / seg_tc(N)) * sl.seg_g, sl.max_c, sl.count);
    k_col_segments<N, M, E><<<g, seg_tc(N) * (M / E), 0, st>>>(sl.descs, sl.tw);
    return hipSuccess;
  }
}

template <int N>
static hipError_t launch_step_n(const StepLaunch& sl, hipStream_t st) {
  hipError_t e;
  if ((e = record(sl.ev[0], st)) != hipSuccess) return e;
  if ((e = launch_rows_e<N>(sl, st, true)) != hipSuccess) return e;
  if ((e = record(sl.ev[1], st)) != hipSuccess) return e;
  if (sl.seg_g > 0) e = launch_col_segments<N>(sl, st);
  if (e == hipSuccess && (sl.seg_g == 0 || sl.seg_mixed)) {
    const int tc = sl.col_tc ? sl.col_tc : DefaultGeom<N>::TC;
    switch (sl.col_e ? sl.col_e : DefaultGeom<N>::COL_E) {
      case 8: e = launch_col_tc<N, 8>(sl, tc, st); break;
      case 16: e = launch_col_tc<N, 16>(sl, tc, st); break;
      default: e = hipErrorInvalidValue;
    }
  }
  if (e != hipSuccess) return e;
  if ((e = record(sl.ev[2], st)) != hipSuccess) return e;
  if ((e = launch_rows_e<N>(sl, st, false)) != hipSuccess) return e;
  if ((e = record(sl.ev[3], st)) != hipSuccess) return e;
  if (!sl.skip_diffusion) {
    const int th = sl.diff_th ? sl.diff_th : DefaultGeom<N>::DIFF_TH;
    const int tiles_x = (N + kDiffTW - 1) / kDiffTW, tiles_y = (N + th - 1) / th;
    const int waves = kDiffThreads / 64;
    const dim3 g_dif((tiles_x * tiles_y + waves - 1) / waves, sl.max_c, sl.count);
    switch (th) {
      case 8: k_diffusion<8><<<g_dif, kDiffThreads, 0, st>>>(sl.descs, sl.sin, sl.sout, N, sl.render, sl.islot); break;
      case 16: k_diffusion<16><<<g_dif, kDiffThreads, 0, st>>>(sl.descs, sl.sin, sl.sout, N, sl.render, sl.islot); break;
      case 32: k_diffusion<32><<<g_dif, kDiffThreads, 0, st>>>(sl.descs, sl.sin, sl.sout, N, sl.render, sl.islot); break;
      default: return hipErrorInvalidValue;
    }
  }
  if ((e = record(sl.ev[4], st)) != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t launch_step(int N, const StepLaunch& sl, hipStream_t st) {
  switch (N) {
    case 16: return launch_step_n<16>(sl, st);
    case 32: return launch_step_n<32>(sl, st);
    case 64: return launch_step_n<64>(sl, st);
    case 128: return launch_step_n<128>(sl, st);
    case 256: return launch_step_n<256>(sl, st);
    case 512: return launch_step_n<512>(sl, st);
    case 1024: return launch_step_n<1024>(sl, st);
    case 2048: return launch_step_n<2048>(sl, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_mass(const CreatureDev* descs, int count, int slot, int max_c, double2* partial, int nblk,
                       hipStream_t st) {
  k_mass<<<dim3(nblk, max_c, count), kMassThreads, 0, st>>>(descs, slot, partial, nblk);
  return hipGetLastError();
}

}  // namespace cn
