// ORACLE - test infrastructure only.  Nothing in the product links, loads or calls this file;
// only tests/, the smoke check and the benchmark's CPU-baseline leg use it, as the checker.
//
// CPU restatement of the reference's CPU implementation `ml-rs-example` (ml-rs/example.rs),
// generalised from its compile-time N=1024, C=3, K=9, TEMP=1 to run-time N, C, K, src/dst maps
// and temp, and templated on the scalar type:
//   cpu_ref<float>  - the reference's arithmetic (radix-2 DIT with a per-butterfly f32 cos/sin
//                     of an f32 angle, /n per inverse pass, source FFT redone per kernel);
//                     this is the "ml-rs-example" CPU baseline;
//   cpu_ref<double> - the same algorithm in FP64 (FP64 angles and trig): the FP64 oracle.
// Built with -O3 -ffp-contract=off and no fast-math: Rust never contracts a*b+c into an FMA.
//
// Parity pins (tests/test_oracle.py): brute-force direct circular convolution and direct DFT
// in FP64 at small N, numpy.fft, the growth closed form, the temp=0 box-mean and uniform
// fixed-point cases of the diffusion, and the mass identity.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

template <typename T>
struct Cpx {
  T re, im;
};
template <typename T>
inline Cpx<T> cadd(Cpx<T> a, Cpx<T> b) { return {a.re + b.re, a.im + b.im}; }  // example.rs:12-14
template <typename T>
inline Cpx<T> csub(Cpx<T> a, Cpx<T> b) { return {a.re - b.re, a.im - b.im}; }  // example.rs:18-20
template <typename T>
inline Cpx<T> cmul(Cpx<T> a, Cpx<T> b) {                                         // example.rs:24-26
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
inline Cpx<T> cfrom_polar(T r, T theta) { return {r * std::cos(theta), r * std::sin(theta)}; }  // :30-32

template <typename T>
constexpr T kPi = (T)3.14159265358979323846264338327950288;

// example.rs:42-71 - in-place forward radix-2 DIT FFT, bit-reversal first.
template <typename T>
void fft_1d(Cpx<T>* data, size_t n) {
  unsigned bits = 0;
  while ((size_t(1) << bits) < n) ++bits;
  for (size_t i = 0; i < n; ++i) {
    size_t j = 0;
    for (unsigned b = 0; b < bits; ++b)
      if (i & (size_t(1) << b)) j |= size_t(1) << (bits - 1 - b);
    if (i < j) std::swap(data[i], data[j]);
  }
  size_t len = 1;
  while (len < n) {
    const size_t half = len;
    len *= 2;
    const T angle = -kPi<T> / (T)half;
    for (size_t i = 0; i < n; i += len)
      for (size_t j = 0; j < half; ++j) {
        const Cpx<T> w = cfrom_polar((T)1, angle * (T)j);
        const Cpx<T> u = data[i + j];
        const Cpx<T> v = cmul(data[i + j + half], w);
        data[i + j] = cadd(u, v);
        data[i + j + half] = csub(u, v);
      }
  }
}

// example.rs:76-88 - conjugate, forward FFT, conjugate and divide by n.
template <typename T>
void ifft_1d(Cpx<T>* data, size_t n) {
  for (size_t i = 0; i < n; ++i) data[i].im = -data[i].im;
  fft_1d(data, n);
  const T nn = (T)n;
  for (size_t i = 0; i < n; ++i) {
    data[i].im = -data[i].im;
    data[i].re /= nn;
    data[i].im /= nn;
  }
}

// example.rs:95-109 / 112-126 - rows, then columns through a gather buffer.
template <typename T, bool INV>
void fft_2d(Cpx<T>* g, size_t n) {
  for (size_t r = 0; r < n; ++r) INV ? ifft_1d(g + r * n, n) : fft_1d(g + r * n, n);
  std::vector<Cpx<T>> col(n);
  for (size_t c = 0; c < n; ++c) {
    for (size_t r = 0; r < n; ++r) col[r] = g[r * n + c];
    INV ? ifft_1d(col.data(), n) : fft_1d(col.data(), n);
    for (size_t r = 0; r < n; ++r) g[r * n + c] = col[r];
  }
}

// example.rs:179-199 - convolution of every kernel (source FFT redone per kernel).  With
// threads > 1 the kernels are shared out over threads; each writes only its own buffer, so the
// results are the same bits as with one thread (the reference runs one).
template <typename T>
void conv_all(size_t n, size_t K, const T* channels, const T* kfft, const uint32_t* src, Cpx<T>* conv,
              int threads = 1) {
  const size_t P = n * n;
  auto one = [&](size_t k) {
    Cpx<T>* g = conv + k * P;
    const T* ch = channels + src[k] * P;
    for (size_t i = 0; i < P; ++i) g[i] = {ch[i], (T)0};
    fft_2d<T, false>(g, n);
    const T* kf = kfft + k * P * 2;
    for (size_t i = 0; i < P; ++i) g[i] = cmul(g[i], Cpx<T>{kf[2 * i], kf[2 * i + 1]});
    fft_2d<T, true>(g, n);
  };
  if (threads <= 1) {
    for (size_t k = 0; k < K; ++k) one(k);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < threads; ++t)
    th.emplace_back([&, t] {
      for (size_t k = (size_t)t; k < K; k += (size_t)threads) one(k);
    });
  for (auto& x : th) x.join();
}

// example.rs:228-266 - bump growth, weighted sum per output channel, A = exp(temp * wsum).
template <typename T>
void growth_phase(size_t n, size_t K, const Cpx<T>* conv, const uint32_t* dst, const T* mu, const T* sigma,
                  const T* w, T temp, uint32_t out_ch, T* affinity) {
  const size_t P = n * n;
  std::vector<T> wsum(P, (T)0);
  for (size_t k = 0; k < K; ++k) {
    if (dst[k] != out_ch) continue;
    const T m = mu[k], s = sigma[k], wk = w[k];
    const T inv_s2 = (T)1 / (s * s);
    for (size_t i = 0; i < P; ++i) {
      const T u = conv[k * P + i].re;
      const T d = u - m;
      const T g = (T)2 * std::exp((T)-0.5 * d * d * inv_s2) - (T)1;
      wsum[i] += g * wk;
    }
  }
  for (size_t i = 0; i < P; ++i) affinity[i] = std::exp(temp * wsum[i]);
}

// example.rs:272-303 - Z = 3x3 periodic sum of A; out = A * sum_3x3(channel / Z).
template <typename T>
void diffusion_phase(size_t n, const T* affinity, const T* channel, T* out) {
  std::vector<T> z(n * n);
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      T s = 0;
      for (size_t di = 0; di < 3; ++di)
        for (size_t dj = 0; dj < 3; ++dj) s += affinity[((i + di + n - 1) % n) * n + (j + dj + n - 1) % n];
      z[i * n + j] = s;
    }
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      T s = 0;
      for (size_t di = 0; di < 3; ++di)
        for (size_t dj = 0; dj < 3; ++dj) {
          const size_t q = ((i + di + n - 1) % n) * n + (j + dj + n - 1) % n;
          s += channel[q] / z[q];
        }
      out[i * n + j] = affinity[i * n + j] * s;
    }
}

// example.rs:164-222 - one synchronous step.
template <typename T>
void step_ml(size_t n, size_t C, size_t K, T* channels, const T* kfft, const uint32_t* src, const uint32_t* dst,
             const T* mu, const T* sigma, const T* w, T temp, int threads = 1) {
  const size_t P = n * n;
  std::vector<Cpx<T>> conv(K * P);
  conv_all(n, K, channels, kfft, src, conv.data(), threads);
  std::vector<T> aff(P), tmp(P);
  for (size_t c = 0; c < C; ++c) {
    growth_phase(n, K, conv.data(), dst, mu, sigma, w, temp, (uint32_t)c, aff.data());
    std::memcpy(tmp.data(), channels + c * P, P * sizeof(T));
    diffusion_phase(n, aff.data(), tmp.data(), channels + c * P);
  }
}

template <typename T>
std::vector<T> widen(const float* p, size_t n) {
  std::vector<T> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = (T)p[i];
  return v;
}

}  // namespace

extern "C" {

// `steps` synchronous steps in place on channels [C][n][n] (f32).  kfft: [K][n][n] interleaved.
void oracle_run_f32(int n, int C, int K, float* channels, const float* kfft, const uint32_t* src,
                    const uint32_t* dst, const float* mu, const float* sigma, const float* w, float temp, int steps) {
  for (int s = 0; s < steps; ++s) step_ml<float>(n, C, K, channels, kfft, src, dst, mu, sigma, w, temp);
}

// oracle_run_f32 with the kernels' convolutions shared out over `threads` threads (the same
// results; for tests that run many full-size steps).
void oracle_run_f32_threads(int n, int C, int K, float* channels, const float* kfft, const uint32_t* src,
                            const uint32_t* dst, const float* mu, const float* sigma, const float* w, float temp,
                            int steps, int threads) {
  for (int s = 0; s < steps; ++s) step_ml<float>(n, C, K, channels, kfft, src, dst, mu, sigma, w, temp, threads);
}

// Same in FP64 (kernel spectra and parameters widened from f32).
void oracle_run_f64(int n, int C, int K, double* channels, const float* kfft, const uint32_t* src,
                    const uint32_t* dst, const float* mu, const float* sigma, const float* w, float temp, int steps) {
  const size_t P = (size_t)n * n;
  auto kf = widen<double>(kfft, (size_t)K * P * 2);
  auto m = widen<double>(mu, K), s = widen<double>(sigma, K), ww = widen<double>(w, K);
  for (int st = 0; st < steps; ++st)
    step_ml<double>(n, C, K, channels, kf.data(), src, dst, m.data(), s.data(), ww.data(), (double)temp);
}

// Real part of every kernel's (normalised) convolution, FP64: out [K][n][n].
void oracle_conv_f64(int n, int K, const double* channels, const float* kfft, const uint32_t* src, double* out) {
  const size_t P = (size_t)n * n;
  auto kf = widen<double>(kfft, (size_t)K * P * 2);
  std::vector<Cpx<double>> conv((size_t)K * P);
  conv_all<double>(n, K, channels, kf.data(), src, conv.data());
  for (size_t i = 0; i < (size_t)K * P; ++i) out[i] = conv[i].re;
}

// Same with the reference's f32 arithmetic.
void oracle_conv_f32(int n, int K, const float* channels, const float* kfft, const uint32_t* src, float* out) {
  const size_t P = (size_t)n * n;
  std::vector<Cpx<float>> conv((size_t)K * P);
  conv_all<float>(n, K, channels, kfft, src, conv.data());
  for (size_t i = 0; i < (size_t)K * P; ++i) out[i] = conv[i].re;
}

// Affinity A of output channel out_ch from given convolution results u [K][n][n] (FP64).
void oracle_growth_f64(int n, int K, const double* u, const uint32_t* dst, const float* mu, const float* sigma,
                       const float* w, float temp, int out_ch, double* affinity) {
  const size_t P = (size_t)n * n;
  std::vector<Cpx<double>> conv((size_t)K * P);
  for (size_t i = 0; i < (size_t)K * P; ++i) conv[i] = {u[i], 0.0};
  auto m = widen<double>(mu, K), s = widen<double>(sigma, K), ww = widen<double>(w, K);
  growth_phase<double>(n, K, conv.data(), dst, m.data(), s.data(), ww.data(), (double)temp, (uint32_t)out_ch,
                       affinity);
}

// Diffusion of one channel given its affinity (FP64).
void oracle_diffusion_f64(int n, const double* affinity, const double* channel, double* out) {
  diffusion_phase<double>(n, affinity, channel, out);
}

// The reference CPU example's creature in its own f32 arithmetic (example.rs:310-491, with
// N = n): sigmoid (:310-312), ring kernels (:338-419: f32 distance from the grid centre,
// sigmoid envelope times Gaussian ring, f32 sum in row-major order, division, ifftshift by a
// quadrant swap :136-152, the f32 radix-2 fft_2d above), seed (:427-458) and growth parameters
// (:466-491).  kernels_fft [9][n][n] interleaved, seed [3][n][n], mu/sigma/w [9], c0/c1 [9].
void oracle_example_creature_f32(int n, float* kernels_fft, float* seed, float* mu, float* sigma, float* w,
                                 uint32_t* c0, uint32_t* c1) {
  const size_t N = (size_t)n, P = N * N, h = N / 2;
  const int C = 3, K = 9;
  const float radii[K] = {0.25f, 0.3167f, 0.3833f, 0.45f, 0.5167f, 0.5833f, 0.65f, 0.7167f, 0.85f};
  const float widths[K] = {0.03f, 0.0411f, 0.0522f, 0.0633f, 0.0744f, 0.0856f, 0.0967f, 0.1078f, 0.13f};
  auto sigmoid = [](float x) { return 0.5f * std::tanh(x * 0.5f) + 0.5f; };
  const float mid = (float)(N / 2), global_r = 10.0f;
  std::vector<float> spatial(P);
  std::vector<Cpx<float>> grid(P);
  for (int k = 0; k < K; ++k) {
    for (size_t i = 0; i < N; ++i)
      for (size_t j = 0; j < N; ++j) {
        const float di = (float)i - mid, dj = (float)j - mid;
        const float dist = std::sqrt(di * di + dj * dj);
        const float d_scaled = dist / (global_r * radii[k]);
        const float sig = sigmoid(-(d_scaled - 1.0f) * 10.0f);
        const float diff = d_scaled - 0.5f;
        const float ker_val = std::exp(-(diff * diff) / (2.0f * widths[k] * widths[k]));
        spatial[i * N + j] = sig * ker_val;
      }
    float total = 0.0f;
    for (size_t i = 0; i < P; ++i) total += spatial[i];
    if (total > 0.0f)
      for (size_t i = 0; i < P; ++i) spatial[i] /= total;
    for (size_t i = 0; i < P; ++i) grid[i] = {spatial[i], 0.0f};
    for (size_t i = 0; i < h; ++i)  // ifftshift_2d: swap quadrants 0 <-> 3, then 1 <-> 2
      for (size_t j = 0; j < h; ++j) std::swap(grid[i * N + j], grid[(i + h) * N + j + h]);
    for (size_t i = 0; i < h; ++i)
      for (size_t j = 0; j < h; ++j) std::swap(grid[i * N + j + h], grid[(i + h) * N + j]);
    fft_2d<float, false>(grid.data(), N);
    for (size_t i = 0; i < P; ++i) {
      kernels_fft[((size_t)k * P + i) * 2] = grid[i].re;
      kernels_fft[((size_t)k * P + i) * 2 + 1] = grid[i].im;
    }
  }
  const float blobs[C][3] = {{0.25f, 0.0f, 0.0f}, {0.22f, 0.04f, 0.0f}, {0.28f, 0.0f, 0.04f}};
  const float half = (float)N / 2.0f;
  for (int c = 0; c < C; ++c) {
    const float inv_sigma2 = 1.0f / (2.0f * blobs[c][0] * blobs[c][0]);
    for (size_t i = 0; i < N; ++i)
      for (size_t j = 0; j < N; ++j) {
        const float x = ((float)j - half) / half, y = ((float)i - half) / half;
        const float dx = x - blobs[c][1], dy = y - blobs[c][2];
        seed[c * P + i * N + j] = std::exp(-(dx * dx + dy * dy) * inv_sigma2);
      }
  }
  for (int k = 0; k < K; ++k) {
    const int in = k % C, out = k / C;
    c0[k] = (uint32_t)in;
    c1[k] = (uint32_t)out;
    const float perm = (float)(in * C + out);
    mu[k] = 0.05f + 0.20f * perm / (float)K;
    sigma[k] = 0.03f + 0.12f * perm / (float)K;
    w[k] = in == out ? 0.5f : 0.25f;
  }
}

// 2D forward FFT (FP64 or the reference's f32 arithmetic), interleaved in place.
void oracle_fft2_f64(int n, double* data) { fft_2d<double, false>(reinterpret_cast<Cpx<double>*>(data), n); }
void oracle_fft2_f32(int n, float* data) { fft_2d<float, false>(reinterpret_cast<Cpx<float>*>(data), n); }

}  // extern "C"
