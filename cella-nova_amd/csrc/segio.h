// Loads and stores of the segmented column passes (k_col_segments, k_col_adj_segments).
//
// A thread's register slot i holds segment position t + c_i, with c_i the plan's compile-time
// offset (in_off for the forward transform's inputs, out_off for its outputs and the inverse's
// inputs and outputs).  Each access is a per-thread byte offset plus a compile-time one, so the
// unrolled loops spend no vector instructions on addresses:
//   - spectrum rows base .. base + M - 1: a resource at row `base` when the segment does not
//     wrap around the torus (every segment but the first and the last), else wrapped per slot;
//   - table rows min(m, M - m) of frequency m = t + c_i: with M/2 and every c_i multiples of the
//     thread count T, each slot lies on one side of M/2 for every t (m = M/2 itself has both
//     forms), so the side is known at compile time and the row is t + c_i or
//     (M - c_i - (T - 1)) + (T - 1 - t);
//   - result rows o0 - R + t + c_i: a resource at row o0 - R (the segment's first input row),
//     the store predicated on the row being one of the segment's outputs.
#pragma once
#include <hip/hip_runtime.h>
#include "bufio.h"
#include "fft.h"
#include "types.h"

namespace cn {

template <int N, int M, int E>
struct SegIO {
  static constexpr int HALF = N / 2, W = HALF + 1, MH = M / 2, T = M / E;
  static constexpr uint32_t kRowBytes = (uint32_t)HALF * 8;

  // Slots of segment rows base .. base + M - 1 of a spectrum plane (wrapping around the torus),
  // column x.
  static __device__ __forceinline__ void load(const char* plane, int base, int t, int x, float2* v) {
    if (base >= 0 && base + M <= N) {
      const BufRes r = make_res(plane + (size_t)base * kRowBytes, (uint32_t)M * kRowBytes);
      const uint32_t vb = (uint32_t)(t * HALF + x) * 8u;
      static_for<0, E>([&](auto I) {
        constexpr int c = in_off<M, E, false>(decltype(I)::value);
        v[I] = buf_load_c64(r, vb, (uint32_t)c * kRowBytes);
      });
    } else {
      const BufRes r = make_res(plane, (uint32_t)N * kRowBytes);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int y = (base + in_pos<M, E, false>(t, i)) & (N - 1);
        v[i] = buf_load_c64(r, (uint32_t)(y * HALF + x) * 8u, 0);
      }
    }
  }

  // Table factors of column x for the slots' frequencies (one kernel's table).
  static __device__ __forceinline__ void table(const float* tab, int t, int x, float* out) {
    const BufRes r = make_res(tab, (uint32_t)(MH + 1) * W * 4u);
    const uint32_t hp = (uint32_t)(t * W + x) * 4u, hn = (uint32_t)((T - 1 - t) * W + x) * 4u;
    static_for<0, E>([&](auto I) {
      constexpr int c = out_off<M, E, false>(decltype(I)::value);
      if constexpr (MH % T == 0 && c + T - 1 <= MH) {
        out[I] = buf_load_f32(r, hp, (uint32_t)c * W * 4u);
      } else if constexpr (MH % T == 0 && c >= MH) {
        out[I] = buf_load_f32(r, hn, (uint32_t)(M - c - (T - 1)) * W * 4u);
      } else {
        const int m = t + c, mm = m <= MH ? m : M - m;
        out[I] = buf_load_f32(r, (uint32_t)(mm * W + x) * 4u, 0);
      }
    });
  }

  // The segment's output rows o0 .. o0 + rows - 1 (slots at positions R .. R + rows - 1) of a
  // result plane in the column passes' result layout.
  static __device__ __forceinline__ void store(char* plane, const float2* w, int t, int x, int o0, int R, int rows) {
    if constexpr (!kColTiled) {
      // The resource may start up to R rows before the plane (first segment): only offsets of
      // stored slots, which are output rows inside the plane, are ever written.
      const BufRes r = make_res(plane + (ptrdiff_t)(o0 - R) * (ptrdiff_t)kRowBytes, (uint32_t)M * kRowBytes);
      const uint32_t vb = (uint32_t)(t * HALF + x) * 8u;
      const int tr = t - R;
      static_for<0, E>([&](auto I) {
        constexpr int c = out_off<M, E, true>(decltype(I)::value);
        if ((unsigned)(tr + c) < (unsigned)rows) buf_store_c64(w[I], r, vb, (uint32_t)c * kRowBytes);
      });
    } else {
      const BufRes r = make_res(plane, (uint32_t)N * kRowBytes);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int q = out_pos<M, E, true>(t, i) - R;
        if (q >= 0 && q < rows) buf_store_c64(w[i], r, (uint32_t)res_index(N, o0 + q, x) * 8u, 0);
      }
    }
  }
};

}  // namespace cn
