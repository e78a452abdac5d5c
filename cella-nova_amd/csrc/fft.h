// Block-cooperative FP32 complex FFT building blocks for gfx950 (wave64).
//
// The transforms are self-sorting Stockham FFTs (Govindaraju et al., "High Performance
// Discrete Fourier Transforms on Graphics Processors", SC'08) done as radix-R passes:
// every thread holds E complex values in registers, each pass performs E/R radix-R
// butterflies per thread in registers, and consecutive passes exchange data through LDS.
// The first pass reads its operands from wherever the caller put them (global memory,
// registers) and the last pass leaves its results in registers, so a caller can fuse
// loads, pointwise products and stores into the transform without extra LDS trips.
//
// Index conventions for an L-point transform with radix sequence R_0, R_1, ...:
//   threads T = L/E per transform, thread t does butterflies j = t + b*T (b < E/R_p),
//   pass p (span Ns = R_0*...*R_{p-1}) reads x[j + m*L/R_p] (m < R_p) into v[b*R_p + m],
//   twiddles by W_{Ns*R_p}^{m*(j mod Ns)}, and writes y[(j - j mod Ns)*R_p + j mod Ns + q*Ns].
// The first pass therefore reads positions j + m*L/R_0 and the last pass leaves output
// position j + q*L/R_last in v[b*R_last + q].
//
// The forward plan uses radix E for every pass except a final remainder pass; the inverse
// plan puts the remainder first.  So the forward transform's output register layout is
// exactly the inverse transform's input register layout, which is what lets a fused
// "forward FFT -> pointwise product -> inverse FFT" keep the spectrum in registers.
//
// Twiddles come from a table tw[i] = exp(-2*pi*i*i/TWN) computed in FP64 on the host and
// rounded once to FP32 (an L-point transform uses stride TWN/L into it), followed by per-pass
// copies of the same values laid out for coalesced reads (types.h, tw_pass_offset).
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "types.h"

namespace cn {

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 c_mulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x / 2); }

// exp(x) from the hardware base-2 exponential (v_exp_f32, ~1 ulp) with the argument split
// x*log2(e) = p + e carried in extended precision: p = RN(x*L2E), e = the rounding error of
// that product (FMA) plus x*(log2(e) - L2E); exp(x) = 2^p * 2^e = 2^p * (1 + e ln 2 + O(e^2)),
// |e| <= 2^-24 |p| + 2^-50 |x|, so the linear term is enough.  Six VALU operations, about
// 2 ulp; inputs whose result under/overflows single precision give 0 / inf like expf.
__device__ __forceinline__ float exp_fast(float x) {
  constexpr float L2E = 1.44269502163f;        // RN(log2 e)
  constexpr float L2E_LO = 1.9259629891e-08f;     // log2(e) - L2E
  const float p = x * L2E;
  const float e = fmaf(x, L2E, -p) + x * L2E_LO;
  const float r = __builtin_amdgcn_exp2f(p);
  return fmaf(r, e * 0.693147180560f, r);
}

// cos(2*pi*k/16), sin(2*pi*k/16) for k = 0..3 (enough for every radix-16 sub-twiddle
// after the special angles 0, 1/8, 1/4, 3/8 are handled exactly).
constexpr float kC16_1 = 0.92387953251128675613f;  // cos(pi/8)
constexpr float kS16_1 = 0.38268343236508977173f;  // sin(pi/8)
constexpr float kSqrtHalf = 0.70710678118654752440f;
constexpr float kC32_1 = 0.98078528040323044913f;  // cos(pi/16)
constexpr float kS32_1 = 0.19509032201612826785f;  // sin(pi/16)
constexpr float kC32_3 = 0.83146961230254523708f;  // cos(3pi/16)
constexpr float kS32_3 = 0.55557023301960222474f;  // sin(3pi/16)

// v * W_R^q with W_R = exp(-2*pi*i/R) (forward) or its conjugate (inverse), q < R/2.
// q and R are compile-time constants after unrolling, so the branches fold away.
template <int R, bool INV>
__device__ __forceinline__ float2 mul_wq(float2 v, int q) {
  // Reduce q/R to an angle in units of 1/32 turn: a = 32*q/R (exact for R <= 32).
  const int a = (32 / R) * q;  // 0..15
  float c, s;                   // W = c - i*s for forward
  if (a == 0) return v;
  if (a == 8) return INV ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x);
  if (a == 4) {
    return INV ? make_float2((v.x - v.y) * kSqrtHalf, (v.x + v.y) * kSqrtHalf)
               : make_float2((v.x + v.y) * kSqrtHalf, (v.y - v.x) * kSqrtHalf);
  }
  if (a == 12) {
    return INV ? make_float2(-(v.x + v.y) * kSqrtHalf, (v.x - v.y) * kSqrtHalf)
               : make_float2((v.y - v.x) * kSqrtHalf, -(v.x + v.y) * kSqrtHalf);
  }
  switch (a) {
    case 2: c = kC16_1; s = kS16_1; break;
    case 6: c = kS16_1; s = kC16_1; break;
    case 10: c = -kS16_1; s = kC16_1; break;
    case 14: c = -kC16_1; s = kS16_1; break;
    case 1: c = kC32_1; s = kS32_1; break;
    case 3: c = kC32_3; s = kS32_3; break;
    case 5: c = kS32_3; s = kC32_3; break;
    case 7: c = kS32_1; s = kC32_1; break;
    case 9: c = -kS32_1; s = kC32_1; break;
    case 11: c = -kS32_3; s = kC32_3; break;
    case 13: c = -kC32_3; s = kS32_3; break;
    default: c = -kC32_1; s = kS32_1; break;  // 15
  }
  // forward: v * (c - i s); inverse: v * (c + i s)
  return INV ? make_float2(v.x * c - v.y * s, v.y * c + v.x * s)
             : make_float2(v.x * c + v.y * s, v.y * c - v.x * s);
}

// In-register R-point DFT (R <= 32), natural order in and out, radix-2 decimation in time.
template <int R, bool INV>
struct Dft {
  static __device__ __forceinline__ void run(float2* a) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) {
      e[i] = a[2 * i];
      o[i] = a[2 * i + 1];
    }
    Dft<R / 2, INV>::run(e);
    Dft<R / 2, INV>::run(o);
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
      const float2 t = mul_wq<R, INV>(o[q], q);
      a[q] = c_add(e[q], t);
      a[q + R / 2] = c_sub(e[q], t);
    }
  }
};
template <bool INV>
struct Dft<1, INV> {
  static __device__ __forceinline__ void run(float2*) {}
};
template <bool INV>
struct Dft<2, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    const float2 x = a[0], y = a[1];
    a[0] = c_add(x, y);
    a[1] = c_sub(x, y);
  }
};
template <bool INV>
struct Dft<4, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    const float2 t0 = c_add(a[0], a[2]), t1 = c_sub(a[0], a[2]);
    const float2 t2 = c_add(a[1], a[3]), d = c_sub(a[1], a[3]);
    const float2 t3 = INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
    a[0] = c_add(t0, t2);
    a[2] = c_sub(t0, t2);
    a[1] = c_add(t1, t3);
    a[3] = c_sub(t1, t3);
  }
};

// Pass plan of an L-point transform with E values per thread.
template <int L, int E>
struct Plan {
  static_assert(E >= 2 && (E & (E - 1)) == 0, "E must be a power of two");
  static_assert(L >= E && (L & (L - 1)) == 0, "L must be a power of two >= E");
  static constexpr int T = L / E;
  static constexpr int lgL = ilog2c(L), lgE = ilog2c(E);
  static constexpr int nfull = lgL / lgE;
  static constexpr int rem = 1 << (lgL - nfull * lgE);
  static constexpr int npass = nfull + (rem > 1 ? 1 : 0);
  static constexpr int radix(int p, bool inv) {
    return rem == 1 ? E : (inv ? (p == 0 ? rem : E) : (p == npass - 1 ? rem : E));
  }
  static constexpr int span(int p, bool inv) {
    int s = 1;
    for (int i = 0; i < p; ++i) s *= radix(i, inv);
    return s;
  }
  static constexpr int first_radix(bool inv) { return radix(0, inv); }
  static constexpr int last_radix(bool inv) { return radix(npass - 1, inv); }
};

// LDS addressing of element `pos` of transform `lane` (0 <= lane < NT) in a block that holds
// NT independent transforms.  INTERLEAVED: (pos + pad)*NT + lane (lane-consecutive threads
// touch consecutive addresses; used for column tiles).  Otherwise lane*LP + pos + pad.
// pad = pos/16 inserts one element every 16: the first pass of a transform writes its outputs
// at stride R (16), which without padding puts every lane of a wave on the same LDS banks.
//
// at(a + c) == at(a) + off(c) whenever (a mod 16) + (c mod 16) < 16; callers use this with a
// compile-time c so that the padded address is one per-butterfly base plus the LDS
// instruction's immediate offset (no per-element index arithmetic).
template <int L, int NT, bool INTERLEAVED>
struct LdsMap {
  static constexpr int LP = L + L / 16;
  static constexpr int size = NT * LP;
  static __device__ __forceinline__ int at(int pos, int lane) {
    const int pp = pos + (pos >> 4);
    return INTERLEAVED ? pp * NT + lane : lane * LP + pp;
  }
  static constexpr int off(int c) { return INTERLEAVED ? (c + c / 16) * NT : c + c / 16; }
};

// LDS copy of the twiddles of one plan (L, E, INV), for transforms whose global twiddle loads
// would share the in-order vector-memory counter with the kernel's own loads and stores (a wait
// for a twiddle would also wait for every older prefetch and store).  For each pass p >= 1
// (span NS, radix R, M = NS R) either the direct block D[m-1][k] = W_M^{m k} when
// (R-1) NS <= 256, or two factors with k = 16 k_hi + k_lo: B[m-1][k_lo] = W_M^{m k_lo} followed
// by A[m-1][k_hi] = W_M^{16 m k_hi}, so that W_M^{m k} = A B (one complex product, ~2 ulp).
template <int L, int E, bool INV>
struct TwLds {
  using PL = Plan<L, E>;
  static constexpr int R(int p) { return PL::radix(p, INV); }
  static constexpr int NS(int p) { return PL::span(p, INV); }
  static constexpr bool direct(int p) { return (R(p) - 1) * NS(p) <= 256; }
  static constexpr int pass_size(int p) {
    return direct(p) ? (R(p) - 1) * NS(p) : (R(p) - 1) * (16 + NS(p) / 16);
  }
  static constexpr int off(int p) {
    int o = 0;
    for (int q = 1; q < p; ++q) o += pass_size(q);
    return o;
  }
  static constexpr int size = off(PL::npass);

  // Index in the global table (layout in types.h) of LDS entry j, 0 <= j < size.
  static __device__ __forceinline__ int src(int j) { return src_from<1>(j); }
  template <int P>
  static __device__ __forceinline__ int src_from(int j) {
    if constexpr (P >= PL::npass) {
      return 0;
    } else {
      constexpr int r = R(P), ns = NS(P), m_n = r * ns;
      if (j < off(P) + pass_size(P)) {
        const int i = j - off(P);
        if constexpr (direct(P)) {
          return tw_pass_offset(ns, r) + i;
        } else {
          constexpr int nh = ns / 16;
          if (i < (r - 1) * 16) return tw_pass_offset(ns, r) + (i / 16) * ns + (i % 16);
          const int ii = i - (r - 1) * 16, m = ii / nh + 1, kh = ii % nh;
          return ((16 * m * kh) * (kTwN / m_n)) % kTwN;
        }
      }
      return src_from<P + 1>(j);
    }
  }
};

// An empty LDS twiddle table (for TwFill with a single table).
struct TwNone {
  static constexpr int size = 0;
  static __device__ __forceinline__ int src(int) { return 0; }
};

// Copies the LDS twiddle tables A then B (each with `size` entries and a `src` map into the
// global table) into one LDS array with all of a thread's global loads issued before any LDS
// store: one memory round trip per block instead of one per table part.  load() can be issued
// before the kernel's other loads complete; store() waits for its own loads only, and the
// caller's next barrier orders the stores before any use.  NTH: the block size.
template <class A, class B, int NTH>
struct TwFill {
  static constexpr int S = A::size + B::size;
  static constexpr int IT = S > 0 ? (S + NTH - 1) / NTH : 1;
  float2 v[IT];
  __device__ __forceinline__ void load(const float2* __restrict__ tw, int tid) {
    gptr<const float2> g = to_global(tw);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      int j = tid + it * NTH;
      j = j < S ? j : 0;
      v[it] = g[j < A::size ? A::src(j) : B::src(j - A::size)];
    }
  }
  __device__ __forceinline__ void store(float2* dst, int tid) const {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int j = tid + it * NTH;
      if (j < S) dst[j] = v[it];
    }
  }
};

// The twiddles of a plan's last pass when it uses two LDS factors (W = A B, see TwLds), formed
// once per thread and kept in registers: a kernel that runs the same plan for many rows or
// kernels, and has registers to spare, saves the two LDS reads and the complex product of
// every such twiddle on every transform (the values are the same products, bit for bit).
template <int L, int E, bool INV>
struct TwRegs {
  using PL = Plan<L, E>;
  using TL = TwLds<L, E, INV>;
  static constexpr int P = PL::npass - 1;
  static constexpr bool kOn = P >= 1 && !TL::direct(P);
  static constexpr int R = PL::radix(P, INV), NS = PL::span(P, INV), NB = E / R;
  float2 w[kOn ? NB * (R - 1) : 1];
  // twl: the TwLds<L, E, INV> block in LDS (filled and ordered by a barrier before the call).
  __device__ __forceinline__ void init(const float2* twl, int t) {
    if constexpr (kOn) {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int k = (t + b * PL::T) & (NS - 1);
        const float2* tb = twl + TL::off(P) + (k & 15);
        const float2* ta = twl + TL::off(P) + (R - 1) * 16 + (k >> 4);
#pragma unroll
        for (int m = 1; m < R; ++m) w[b * (R - 1) + m - 1] = c_mul(ta[(m - 1) * (NS / 16)], tb[(m - 1) * 16]);
      }
    }
  }
};

// One Stockham pass on register data (twiddle + radix-R butterflies) for thread t.  Pass P of
// the plan; LT: the twiddles come from TwLds<L, E, INV> at `tw` (LDS), else from the global
// table; REG: the last pass's two-factor twiddles come from a TwRegs at `wr`.
template <int L, int E, int R, int NS, bool INV, int TWN, int P, bool LT, bool REG = false>
__device__ __forceinline__ void pass_compute(float2* v, int t, const float2* __restrict__ tw,
                                             const float2* wr = nullptr) {
  constexpr int T = L / E;
  constexpr int NB = E / R;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
    if constexpr (NS > 1) {
      const int k = j & (NS - 1);
      if constexpr (LT) {
        using TL = TwLds<L, E, INV>;
        static_assert(TL::R(P) == R && TL::NS(P) == NS, "LDS twiddle layout / plan mismatch");
        if constexpr (TL::direct(P)) {
          const float2* twp = tw + TL::off(P) + k;
#pragma unroll
          for (int m = 1; m < R; ++m) {
            const float2 w = twp[(m - 1) * NS];
            v[b * R + m] = INV ? c_mulc(v[b * R + m], w) : c_mul(v[b * R + m], w);
          }
        } else if constexpr (REG && P == TwRegs<L, E, INV>::P) {
#pragma unroll
          for (int m = 1; m < R; ++m) {
            const float2 w = wr[b * (R - 1) + m - 1];
            v[b * R + m] = INV ? c_mulc(v[b * R + m], w) : c_mul(v[b * R + m], w);
          }
        } else {
          const float2* tb = tw + TL::off(P) + (k & 15);
          const float2* ta = tw + TL::off(P) + (R - 1) * 16 + (k >> 4);
#pragma unroll
          for (int m = 1; m < R; ++m) {
            const float2 w = c_mul(ta[(m - 1) * (NS / 16)], tb[(m - 1) * 16]);
            v[b * R + m] = INV ? c_mulc(v[b * R + m], w) : c_mul(v[b * R + m], w);
          }
        }
      } else {
        // Per-pass block of W_{NS*R}^{m*k}, k fastest: a wavefront's lanes read adjacent entries.
        constexpr int off = tw_pass_offset(NS, R);
        static_assert(off >= 0 && TWN == kTwN, "pass shape outside the twiddle table");
        const float2* twp = tw + off + k;
#pragma unroll
        for (int m = 1; m < R; ++m) {
          const float2 w = twp[(m - 1) * NS];
          v[b * R + m] = INV ? c_mulc(v[b * R + m], w) : c_mul(v[b * R + m], w);
        }
      }
    }
    Dft<R, INV>::run(v + b * R);
  }
}

// Orders the LDS exchange of a transform.  WAVE: all threads of each transform belong to one
// wavefront, and a wavefront performs its LDS accesses in program order, so the compiler only
// has to keep the accesses in order (the __syncwarp sequence); otherwise a block barrier.
template <bool WAVE>
__device__ __forceinline__ void xsync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Row-layout transforms (thread = lane*T + t) whose T threads divide a wavefront.
template <int L, int E>
constexpr bool wave_rows() {
  return 64 % (L / E) == 0;
}

// Hides the value of a lane index from the optimiser.  Index arithmetic derived from a
// laundered value is recomputed where it is used instead of being hoisted out of the
// caller's loops as dozens of loop-invariant 64-bit addresses (which spill).
__device__ __forceinline__ int launder(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Runs passes P..npass-1 of an L-point transform (LT: twiddles from an LDS TwLds block).  On
// entry v holds the inputs of pass P (layout of pass P); on exit it holds the outputs of the last
// pass.  `lds` must hold
// LdsMap<L,NT,IL>::size elements; every exchange is bracketed by barriers (wavefront-level for
// row-layout transforms that fit in a wavefront, block-level otherwise), so the routine can be
// called repeatedly on the same buffer.
template <int L, int E, int NT, bool IL, bool INV, int TWN, int P = 0, bool LT = false, bool REG = false>
__device__ __forceinline__ void fft_run(float2* v, float2* lds, int t, int lane,
                                        const float2* __restrict__ tw, const float2* wr = nullptr) {
  using PL = Plan<L, E>;
  constexpr int R = PL::radix(P, INV);
  constexpr int NS = PL::span(P, INV);
  t = launder(t);
  pass_compute<L, E, R, NS, INV, TWN, P, LT, REG>(v, t, tw, wr);
  if constexpr (P + 1 < PL::npass) {
    constexpr int T = L / E;
    constexpr int R2 = PL::radix(P + 1, INV);
    using LM = LdsMap<L, NT, IL>;
    // Writes go to base + q*NS.  base mod 16 is k < NS when NS*R >= 16 (base - k is a multiple
    // of NS*R), a multiple of R when NS == 1, and q*NS is a multiple of 16 when NS >= 16: in
    // each case the padded index splits into at(base) + off(q*NS).  Reads go to j + m*L/R2.
    constexpr bool kSplitW = (NS % 16 == 0) || (NS == 1 && 16 % R == 0) || (16 % NS == 0 && NS * R >= 16);
    constexpr bool kSplitR = (L / R2) % 16 == 0;
    constexpr bool kWave = !IL && wave_rows<L, E>();
    xsync<kWave>();
#pragma unroll
    for (int b = 0; b < E / R; ++b) {
      const int j = t + b * T;
      const int k = j & (NS - 1);
      const int base = (j - k) * R + k;
      if constexpr (kSplitW) {
        float2* p = lds + LM::at(base, lane);
#pragma unroll
        for (int q = 0; q < R; ++q) p[LM::off(q * NS)] = v[b * R + q];
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) lds[LM::at(base + q * NS, lane)] = v[b * R + q];
      }
    }
    xsync<kWave>();
#pragma unroll
    for (int b = 0; b < E / R2; ++b) {
      const int j = t + b * T;
      if constexpr (kSplitR) {
        const float2* p = lds + LM::at(j, lane);
#pragma unroll
        for (int m = 0; m < R2; ++m) v[b * R2 + m] = p[LM::off(m * (L / R2))];
      } else {
#pragma unroll
        for (int m = 0; m < R2; ++m) v[b * R2 + m] = lds[LM::at(j + m * (L / R2), lane)];
      }
    }
    fft_run<L, E, NT, IL, INV, TWN, P + 1, LT, REG>(v, lds, t, lane, tw, wr);
  }
}

// Position (within the transform) of register slot i for the first pass's inputs and the
// last pass's outputs: thread t holds position t + in_off(i) / t + out_off(i).
template <int L, int E, bool INV>
__host__ __device__ constexpr int in_off(int i) {
  using PL = Plan<L, E>;
  return (i / PL::first_radix(INV)) * PL::T + (i % PL::first_radix(INV)) * (L / PL::first_radix(INV));
}
template <int L, int E, bool INV>
__host__ __device__ constexpr int out_off(int i) {
  using PL = Plan<L, E>;
  return (i / PL::last_radix(INV)) * PL::T + (i % PL::last_radix(INV)) * (L / PL::last_radix(INV));
}
template <int L, int E, bool INV>
__device__ __forceinline__ int in_pos(int t, int i) {
  using PL = Plan<L, E>;
  constexpr int R = PL::first_radix(INV);
  return t + (i / R) * PL::T + (i % R) * (L / R);
}
template <int L, int E, bool INV>
__device__ __forceinline__ int out_pos(int t, int i) {
  using PL = Plan<L, E>;
  constexpr int R = PL::last_radix(INV);
  return t + (i / R) * PL::T + (i % R) * (L / R);
}

// f(std::integral_constant<int, I>{}) for I = 0 .. n-1: a loop whose index is a constant
// expression in the body (compile-time choices per register slot).
template <int I, int n, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < n) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, n>(f);
  }
}

}  // namespace cn
