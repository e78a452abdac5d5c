// Real-row transforms built on the block FFT (fft.h): an N-point real row is handled as an
// N/2-point complex transform of z[n] = x[2n] + i x[2n+1] plus a split/merge step.  The half
// spectrum of a row is stored as N/2 complex values with slot 0 packing the two real bins
// (X[0], X[N/2]) as (re, im).
#pragma once
#include "bufio.h"
#include "fft.h"
#include "types.h"

namespace cn {

// W_N^p, 0 <= p < N/2, for the real-row split/merge from a small LDS table pair:
// p = 32 p_hi + p_lo, lo[p_lo] = W_N^{p_lo}, hi[p_hi] = W_N^{32 p_hi}, W_N^p = hi * lo.
template <int N>
struct SplitTw {
  static constexpr int NLO = 32, NHI = (N / 2 + 31) / 32;
  static constexpr int size = NLO + NHI;
  static __device__ __forceinline__ int src(int i) { return ((i < NLO ? i : 32 * (i - NLO)) * (kTwN / N)) % kTwN; }
  static __device__ __forceinline__ float2 get(const float2* d, int p) {
    return c_mul(d[NLO + (p >> 5)], d[p & 31]);
  }
};

// After the forward L-point transform of z (L = N/2), v holds Z at out_pos positions; turn it
// into the real row's half spectrum and store it to `out` (L complex).  Uses `lds` (the same
// buffer as the transform) for the Z[L-k] partner exchange.
// LT: `tw` is a SplitTw<N> table in LDS, else the global table.
template <int N, int E, int ROWS, bool LT = false>
__device__ __forceinline__ void r2c_finish(const float2* v, float2* lds, int t, int lane,
                                           const float2* __restrict__ tw, gptr<float2> out, bool store) {
  constexpr int L = N / 2;
  using LM = LdsMap<L, ROWS, false>;
  constexpr bool kWave = wave_rows<L, E>();
  xsync<kWave>();
  if constexpr ((L / E) % 16 == 0 && (L / Plan<L, E>::last_radix(false)) % 16 == 0) {
    // out_pos(t, i) = t + c_i with c_i a multiple of 16: base + immediate offsets.
    float2* p = lds + LM::at(t, lane);
#pragma unroll
    for (int i = 0; i < E; ++i) p[LM::off(out_pos<L, E, false>(0, i))] = v[i];
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) lds[LM::at(out_pos<L, E, false>(t, i), lane)] = v[i];
  }
  xsync<kWave>();
  // Branch-free: every lane loads its partner and twiddle (p = 0 reads itself and tw[0]) and
  // the packed (X[0], X[N/2]) slot is selected at the end, so no load sits in a divergent
  // branch with its own wait.
  float2 w[E], zq[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int p = out_pos<L, E, false>(t, i);
    w[i] = LT ? SplitTw<N>::get(tw, p) : tw[p * (kTwN / N)];
    zq[i] = lds[LdsMap<L, ROWS, false>::at((L - p) & (L - 1), lane)];
  }
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int p = out_pos<L, E, false>(t, i);
    const float2 zp = v[i];
    const float2 ze = make_float2(0.5f * (zp.x + zq[i].x), 0.5f * (zp.y - zq[i].y));
    const float2 dd = make_float2(zp.x - zq[i].x, zp.y + zq[i].y);  // zp - conj(zq)
    const float2 zo = make_float2(0.5f * dd.y, -0.5f * dd.x);        // dd / (2i)
    const float2 X = c_add(ze, c_mul(w[i], zo));
    const float2 X0 = make_float2(zp.x + zp.y, zp.x - zp.y);
    if (store) out[p] = (p == 0) ? X0 : X;
  }
}

// Loads a half-spectrum row X (L = N/2 complex) into the pairs (X[p], X[L-p]) that the inverse
// transform's first pass needs; c2r_form turns them into the L-point input Z2 of the inverse
// transform whose unnormalised output is z[n] = x[2n] + i x[2n+1], x being the unnormalised
// N-point real inverse of X.  Split in two so that callers can issue the loads of the next
// row early (software prefetch).
template <int N, int E>
__device__ __forceinline__ void c2r_load(gptr<const float2> X, int t, float2* xp, float2* xq) {
  constexpr int L = N / 2;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int p = in_pos<L, E, true>(t, i);
    xp[i] = X[p];
    xq[i] = X[(L - p) & (L - 1)];
  }
}

template <int N, int E, bool LT = false>
__device__ __forceinline__ void c2r_form(const float2* xp, const float2* xq, int t, const float2* __restrict__ tw,
                                         float2* v) {
  constexpr int L = N / 2;
  // Branch-free (see r2c_finish): the twiddle is loaded for every lane, slot 0 selected last.
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int p = in_pos<L, E, true>(t, i);
    const float2 e2 = make_float2(xp[i].x + xq[i].x, xp[i].y - xq[i].y);  // xp + conj(xq)
    const float2 dd = make_float2(xp[i].x - xq[i].x, xp[i].y + xq[i].y);  // xp - conj(xq)
    const float2 o2 = c_mulc(dd, LT ? SplitTw<N>::get(tw, p) : tw[p * (kTwN / N)]);  // * W_N^{-p}
    const float2 g = make_float2(e2.x - o2.y, e2.y + o2.x);               // e2 + i*o2
    const float2 g0 = make_float2(xp[i].x + xp[i].y, xp[i].x - xp[i].y);
    v[i] = (p == 0) ? g0 : g;
  }
}

// Each lane loads only its own X[p] and takes the partner X[L-p] from the other lanes through
// `lds` (the transform's exchange buffer), ordered at wavefront level for single-wavefront rows
// (wave_rows<N/2, E>()) and by block barriers otherwise.  Halves the row's global loads (and
// the registers of a prefetched row) against c2r_load.
template <int N, int E>
__device__ __forceinline__ void c2r_load_own(gptr<const float2> X, int t, float2* xp) {
  constexpr int L = N / 2;
#pragma unroll
  for (int i = 0; i < E; ++i) xp[i] = X[in_pos<L, E, true>(t, i)];
}

// c2r_load_own for row y of a column-pass result plane (res_index).  Column-tiled: when the thread
// count per row is a multiple of 4, element i of thread t is column t + c_i with c_i a multiple
// of 4, so the tile (t >> 2) + c_i / 4 and the column inside it (t & 3) split into a per-thread
// offset and a per-element constant.
template <int N, int E>
__device__ __forceinline__ void c2r_load_own_ct(const void* plane, int y, int t, float2* xp) {
  constexpr int L = N / 2, T = L / (E < L ? E : L);
  if constexpr (!kColTiled) {
    const BufRes r = make_res(plane, (uint32_t)N * L * 8u);
    const uint32_t vo = (uint32_t)(y * L + t) * 8u;
#pragma unroll
    for (int i = 0; i < E; ++i) xp[i] = buf_load_c64(r, vo, (uint32_t)in_pos<L, E, true>(0, i) * 8u);
  } else if constexpr (T % 4 == 0) {
    const BufRes r = make_res(plane, (uint32_t)N * L * 8u);
    const uint32_t vo = (uint32_t)(((t >> 2) * N + y) * 4 + (t & 3)) * 8u;
#pragma unroll
    for (int i = 0; i < E; ++i) xp[i] = buf_load_c64(r, vo, (uint32_t)(in_pos<L, E, true>(0, i) >> 2) * N * 32u);
  } else {
    gptr<const float2> X = c64(plane);
#pragma unroll
    for (int i = 0; i < E; ++i) xp[i] = X[res_index(N, y, in_pos<L, E, true>(t, i))];
  }
}

// c2r_start for row y of a column-pass result plane (res_index).
template <int N, int E, bool LT = false>
__device__ __forceinline__ void c2r_start_ct(const void* plane, int y, int t, const float2* __restrict__ tw,
                                             float2* v) {
  constexpr int L = N / 2;
  gptr<const float2> X = c64(plane);
  float2 xp[E], xq[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int p = in_pos<L, E, true>(t, i);
    xp[i] = X[res_index(N, y, p)];
    xq[i] = X[res_index(N, y, (L - p) & (L - 1))];
  }
  c2r_form<N, E, LT>(xp, xq, t, tw, v);
}

template <int N, int E, int ROWS>
__device__ __forceinline__ void c2r_partner(const float2* xp, float2* lds, int t, int lane, float2* xq) {
  constexpr int L = N / 2;
  constexpr bool kWave = wave_rows<L, E>();
  using LM = LdsMap<L, ROWS, false>;
  xsync<kWave>();
#pragma unroll
  for (int i = 0; i < E; ++i) lds[LM::at(in_pos<L, E, true>(t, i), lane)] = xp[i];
  xsync<kWave>();
#pragma unroll
  for (int i = 0; i < E; ++i) xq[i] = lds[LM::at((L - in_pos<L, E, true>(t, i)) & (L - 1), lane)];
}

template <int N, int E, bool LT = false>
__device__ __forceinline__ void c2r_start(gptr<const float2> X, int t, const float2* __restrict__ tw,
                                          float2* v) {
  float2 xp[E], xq[E];
  c2r_load<N, E>(X, t, xp, xq);
  c2r_form<N, E, LT>(xp, xq, t, tw, v);
}

}  // namespace cn
