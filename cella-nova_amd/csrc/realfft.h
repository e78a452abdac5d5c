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

// Split-table twiddle W_N^p for p = t + c with c a multiple of 32: the low factor depends on t
// alone (`lo`, loaded once per thread) and the high one sits at a compile-time offset.
template <int N, int C>
__device__ __forceinline__ float2 split_tw(const float2* tw, float2 lo, int t) {
  static_assert(C % 32 == 0, "split_tw needs offsets that are multiples of 32");
  return c_mul((tw + SplitTw<N>::NLO + (t >> 5))[C / 32], lo);
}

// LDS element of the partner position (L - p) & (L - 1), p = t + c, of row `lane` in a row
// exchange buffer.  For 0 < c <= CMAX with L - c and L - CMAX multiples of 16 this is a
// per-thread base (pb, see partner_base) plus a compile-time offset; otherwise the general map.
template <int L, int ROWS, int CMAX>
__device__ __forceinline__ const float2* partner_base(const float2* lds, int t, int lane) {
  using LM = LdsMap<L, ROWS, false>;
  return lds + lane * LM::LP + (L - CMAX) + (L - CMAX) / 16 - t - ((t + 15) >> 4);
}
template <int L, int ROWS, int C, int CMAX>
__device__ __forceinline__ float2 partner(const float2* lds, const float2* pb, int t, int lane) {
  if constexpr (C > 0 && C <= CMAX && (L - C) % 16 == 0 && (L - CMAX) % 16 == 0) {
    return pb[(CMAX - C) + (CMAX - C) / 16];
  } else {
    return lds[LdsMap<L, ROWS, false>::at((L - (t + C)) & (L - 1), lane)];
  }
}

// After the forward L-point transform of z (L = N/2), v holds Z at out_pos positions; turn it
// into the real row's half spectrum and store it to `out` (L complex).  Uses `lds` (the same
// buffer as the transform) for the Z[L-k] partner exchange.
// LT: `tw` is a SplitTw<N> table in LDS, else the global table.
// Register slot i holds position p = t + c_i with c_i a compile-time constant: partner and
// twiddle addresses are per-thread bases plus immediate offsets where the offsets allow, and
// only the slot with c_i = 0 can hold the packed (X[0], X[N/2]) position p = 0.
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
  constexpr int CMAX = [] {
    int m = 0;
    for (int i = 0; i < E; ++i) m = out_off<L, E, false>(i) > m ? out_off<L, E, false>(i) : m;
    return m;
  }();
  const float2* pb = partner_base<L, ROWS, CMAX>(lds, t, lane);
  const float2 lo = LT ? tw[t & 31] : make_float2(1.f, 0.f);
  // Branch-free: every lane loads its partner and twiddle (p = 0 reads itself and tw[0]) and
  // the packed (X[0], X[N/2]) slot is selected at the end, so no load sits in a divergent
  // branch with its own wait.
  float2 w[E], zq[E];
  static_for<0, E>([&](auto I) {
    constexpr int c = out_off<L, E, false>(decltype(I)::value);
    const int p = t + c;
    if constexpr (LT && c % 32 == 0) w[I] = split_tw<N, c>(tw, lo, t);
    else w[I] = LT ? SplitTw<N>::get(tw, p) : tw[p * (kTwN / N)];
    zq[I] = partner<L, ROWS, c, CMAX>(lds, pb, t, lane);
  });
  static_for<0, E>([&](auto I) {
    constexpr int i = decltype(I)::value, c = out_off<L, E, false>(i);
    const int p = t + c;
    const float2 zp = v[i];
    const float2 ze = make_float2(0.5f * (zp.x + zq[i].x), 0.5f * (zp.y - zq[i].y));
    const float2 dd = make_float2(zp.x - zq[i].x, zp.y + zq[i].y);  // zp - conj(zq)
    const float2 zo = make_float2(0.5f * dd.y, -0.5f * dd.x);        // dd / (2i)
    const float2 X = c_add(ze, c_mul(w[i], zo));
    if constexpr (c == 0) {
      const float2 X0 = make_float2(zp.x + zp.y, zp.x - zp.y);
      if (store) out[p] = (p == 0) ? X0 : X;
    } else {
      if (store) out[p] = X;
    }
  });
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

// dc is subtracted from the row's DC bin, so that every real output comes out less dc.
template <int N, int E, bool LT = false>
__device__ __forceinline__ void c2r_form(const float2* xp, const float2* xq, int t, const float2* __restrict__ tw,
                                         float2* v, float dc = 0.f) {
  constexpr int L = N / 2;
  const float2 lo = LT ? tw[t & 31] : make_float2(1.f, 0.f);
  // Branch-free (see r2c_finish): the twiddle is loaded for every lane; only the slot with a
  // zero offset can hold p = 0, whose packed pair is selected at the end.
  static_for<0, E>([&](auto I) {
    constexpr int i = decltype(I)::value, c = in_off<L, E, true>(i);
    const int p = t + c;
    float2 W;
    if constexpr (LT && c % 32 == 0) W = split_tw<N, c>(tw, lo, t);
    else W = LT ? SplitTw<N>::get(tw, p) : tw[p * (kTwN / N)];
    const float2 e2 = make_float2(xp[i].x + xq[i].x, xp[i].y - xq[i].y);  // xp + conj(xq)
    const float2 dd = make_float2(xp[i].x - xq[i].x, xp[i].y + xq[i].y);  // xp - conj(xq)
    const float2 o2 = c_mulc(dd, W);                                       // * W_N^{-p}
    const float2 g = make_float2(e2.x - o2.y, e2.y + o2.x);               // e2 + i*o2
    if constexpr (c == 0) {
      const float2 g0 = make_float2(xp[i].x + xp[i].y - dc, xp[i].x - xp[i].y - dc);
      v[i] = (p == 0) ? g0 : g;
    } else {
      v[i] = g;
    }
  });
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
  constexpr int CMAX = [] {
    int m = 0;
    for (int i = 0; i < E; ++i) m = in_off<L, E, true>(i) > m ? in_off<L, E, true>(i) : m;
    return m;
  }();
  xsync<kWave>();
  if constexpr ((L / E) % 16 == 0 && (L / Plan<L, E>::first_radix(true)) % 16 == 0) {
    float2* pw = lds + LM::at(t, lane);  // in_pos(t, i) = t + c_i with c_i a multiple of 16
#pragma unroll
    for (int i = 0; i < E; ++i) pw[LM::off(in_pos<L, E, true>(0, i))] = xp[i];
  } else {
#pragma unroll
    for (int i = 0; i < E; ++i) lds[LM::at(in_pos<L, E, true>(t, i), lane)] = xp[i];
  }
  xsync<kWave>();
  const float2* pb = partner_base<L, ROWS, CMAX>(lds, t, lane);
  static_for<0, E>([&](auto I) {
    constexpr int c = in_off<L, E, true>(decltype(I)::value);
    xq[I] = partner<L, ROWS, c, CMAX>(lds, pb, t, lane);
  });
}

template <int N, int E, bool LT = false>
__device__ __forceinline__ void c2r_start(gptr<const float2> X, int t, const float2* __restrict__ tw,
                                          float2* v) {
  float2 xp[E], xq[E];
  c2r_load<N, E>(X, t, xp, xq);
  c2r_form<N, E, LT>(xp, xq, t, tw, v);
}

}  // namespace cn
