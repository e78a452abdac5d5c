// Synthetic creature generator (host, FP64): Lenia ring kernels and their spectra, growth
// parameters and seed patch for the benchmark distributions.  The reference's own generators
// (ml-rs/example.rs:338-491, train/generate_kernel_json.py:106-215) build sigmoid-Gaussian
// rings; the benchmark creatures use the classic Lenia shell kernel instead, defined here by
// its closed form.  The output spectrum format is the reference's .bin layout
// (config.rs:49-55): row-major complex64, unnormalised FFT of the origin-centred kernel.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>
#include "common.h"
#include "prng.h"

namespace cn {
namespace {

// In-place radix-2 complex FFT of length n (power of two), FP64, forward (exp(-i...)).
void fft1d(double* re, double* im, size_t n, const std::vector<double>& cs, const std::vector<double>& sn,
           size_t table_n) {
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(re[i], re[j]);
      std::swap(im[i], im[j]);
    }
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const size_t half = len / 2, step = table_n / len;
    for (size_t i = 0; i < n; i += len)
      for (size_t j = 0; j < half; ++j) {
        const double wr = cs[j * step], wi = sn[j * step];
        const double xr = re[i + j + half], xi = im[i + j + half];
        const double vr = xr * wr - xi * wi, vi = xr * wi + xi * wr;
        re[i + j + half] = re[i + j] - vr;
        im[i + j + half] = im[i + j] - vi;
        re[i + j] += vr;
        im[i + j] += vi;
      }
  }
}

void fft2d(std::vector<double>& re, std::vector<double>& im, size_t n) {
  std::vector<double> cs(n / 2 + 1), sn(n / 2 + 1);
  for (size_t i = 0; i <= n / 2; ++i) {
    const double a = -2.0 * M_PI * (double)i / (double)n;
    cs[i] = std::cos(a);
    sn[i] = std::sin(a);
  }
  for (size_t r = 0; r < n; ++r) fft1d(&re[r * n], &im[r * n], n, cs, sn, n);
  std::vector<double> cr(n), ci(n);
  for (size_t c = 0; c < n; ++c) {
    for (size_t r = 0; r < n; ++r) {
      cr[r] = re[r * n + c];
      ci[r] = im[r * n + c];
    }
    fft1d(cr.data(), ci.data(), n, cs, sn, n);
    for (size_t r = 0; r < n; ++r) {
      re[r * n + c] = cr[r];
      im[r * n + c] = ci[r];
    }
  }
}

}  // namespace

// Growth parameters mu ~ U[0.05, 0.15], sigma ~ U[0.03, 0.06] and Dirichlet(1) weights over
// the kernels that share a destination channel (normalised i.i.d. Exp(1) draws).
void generate_growth(const cn_creature_params* p, float* mu, float* sigma, float* weights) {
  const uint32_t K = p->num_kernels;
  const uint64_t sd = p->seed, cr = p->creature;
  std::vector<double> ew(K);
  for (uint32_t k = 0; k < K; ++k) ew[k] = -std::log1p(-uniform01(sd, cr, kWeight, k));  // Exp(1)
  for (uint32_t k = 0; k < K; ++k) {
    double tot = 0.0;
    for (uint32_t j = 0; j < K; ++j)
      if (p->dst[j] == p->dst[k]) tot += ew[j];
    if (mu) mu[k] = (float)(0.05 + 0.10 * uniform01(sd, cr, kMu, k));
    if (sigma) sigma[k] = (float)(0.03 + 0.03 * uniform01(sd, cr, kSigma, k));
    if (weights) weights[k] = (float)(ew[k] / tot);
  }
}

cn_status check_creature_params(const cn_creature_params* p) {
  if (!p) return set_error(CN_ERR_INVALID_ARG, "null params");
  const uint32_t n = p->n, C = p->num_channels, K = p->num_kernels;
  if (!is_pow2(n) || n < 2 || n > 8192) return set_error(CN_ERR_UNSUPPORTED, "n = %u must be a power of two", n);
  if (C < 1 || C > 16 || K < 1 || K > 256) return set_error(CN_ERR_UNSUPPORTED, "C = %u, K = %u out of range", C, K);
  if (!p->src || !p->dst || !p->radius) return set_error(CN_ERR_INVALID_ARG, "null src/dst/radius");
  for (uint32_t k = 0; k < K; ++k) {
    if (p->src[k] >= C || p->dst[k] >= C) return set_error(CN_ERR_INVALID_ARG, "src/dst[%u] out of range", k);
    if (p->radius[k] < 1) return set_error(CN_ERR_INVALID_ARG, "radius[%u] must be >= 1", k);
  }
  return CN_OK;
}

}  // namespace cn

using namespace cn;

extern "C" cn_status cn_generate_creature(const cn_creature_params* p, float* kernels_fft, double* spatial,
                                          float* state, float* mu, float* sigma, float* weights) {
  clear_error();
  const cn_status st = check_creature_params(p);
  if (st != CN_OK) return st;
  const uint32_t n = p->n, C = p->num_channels, K = p->num_kernels;
  const size_t N = n, P = N * N;
  const uint64_t sd = p->seed, cr = p->creature;

  generate_growth(p, mu, sigma, weights);

  // Seed: U[0,1) on the central patch of every channel.
  if (state) {
    const SeedPatch sp = seed_patch_of(n, p->patch);
    std::memset(state, 0, (size_t)C * P * sizeof(float));
    for (uint32_t c = 0; c < C; ++c)
      for (size_t i = 0; i < sp.side; ++i)
        for (size_t j = 0; j < sp.side; ++j)
          state[c * P + (sp.origin + i) * N + (sp.origin + j)] = seed_value(sd, cr, c, (uint32_t)i, (uint32_t)j, sp.side);
  }

  if (!kernels_fft && !spatial) return CN_OK;
  // Kernels: one worker thread per kernel (bounded by the host's hardware threads).
  auto build = [&](uint32_t k) {
    const RingParams rp = ring_params(sd, cr, k, p->radius[k]);
    std::vector<double> re(P, 0.0), im(P, 0.0);
    double tot = 0.0;
    for (size_t y = 0; y < N; ++y) {
      const double dy = wrap_offset((uint32_t)y, n);
      for (size_t x = 0; x < N; ++x) {
        const double v = ring_value(rp, wrap_offset((uint32_t)x, n), dy);
        re[y * N + x] = v;
        tot += v;
      }
    }
    if (tot > 0.0)
      for (size_t i = 0; i < P; ++i) re[i] /= tot;
    if (spatial) std::memcpy(spatial + (size_t)k * P, re.data(), P * sizeof(double));
    if (kernels_fft) {
      fft2d(re, im, N);
      float* o = kernels_fft + (size_t)k * P * 2;
      for (size_t i = 0; i < P; ++i) {
        o[2 * i] = (float)re[i];
        o[2 * i + 1] = (float)im[i];
      }
    }
  };
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  for (uint32_t k0 = 0; k0 < K; k0 += hw) {
    std::vector<std::thread> th;
    for (uint32_t k = k0; k < std::min(K, k0 + hw); ++k) th.emplace_back(build, k);
    for (auto& t : th) t.join();
  }
  return CN_OK;
}

// ---------------------------------------------------------------------------------------
// The creature of the reference's CPU example (ml-rs/example.rs:310-491), at any grid size n.
// ---------------------------------------------------------------------------------------
namespace {

constexpr int kExC = 3, kExK = 9;
// Ring radii and widths of the example's 9 kernels, as fractions of global_r = 10 (:347-353).
constexpr double kExRadius[kExK] = {0.25, 0.3167, 0.3833, 0.45, 0.5167, 0.5833, 0.65, 0.7167, 0.85};
constexpr double kExWidth[kExK] = {0.03, 0.0411, 0.0522, 0.0633, 0.0744, 0.0856, 0.0967, 0.1078, 0.13};
// Seed blobs (sigma, x offset, y offset) per channel (:430).
constexpr double kExBlob[kExC][3] = {{0.25, 0.0, 0.0}, {0.22, 0.04, 0.0}, {0.28, 0.0, 0.04}};

// sigmoid(x) = tanh(x / 2) / 2 + 1/2 (:310-312).
double ex_sigmoid(double x) { return 0.5 * std::tanh(x * 0.5) + 0.5; }

// Kernel k at the grid point (i, j) of the centred layout: a Gaussian ring at half the scaled
// radius under a sigmoid envelope that ends at the scaled radius (:362-383).
double ex_kernel_value(int k, size_t i, size_t j, size_t n) {
  const double mid = (double)(n / 2), di = (double)i - mid, dj = (double)j - mid;
  const double ds = std::sqrt(di * di + dj * dj) / (10.0 * kExRadius[k]);
  const double diff = ds - 0.5, w = kExWidth[k];
  return ex_sigmoid(-(ds - 1.0) * 10.0) * std::exp(-(diff * diff) / (2.0 * w * w));
}

}  // namespace

// Kernels (:338-419): each ring kernel normalised to sum 1 (:390-397), moved so that its centre
// (n/2, n/2) sits at the origin (ifftshift, :136-152, :408: output (y, x) takes the centred value at
// ((y + n/2) mod n, (x + n/2) mod n)), and transformed (:414; FP64 here, rounded to complex64).
// Seed (:427-458): channel c = exp(-((x - ox)^2 + (y - oy)^2) / (2 s^2)) with x = (j - n/2)/(n/2),
// y = (i - n/2)/(n/2).  Growth (:466-491) in the example's f32 arithmetic: kernel k reads channel
// k % 3 and writes channel k / 3, perm = 3 (k % 3) + k / 3, mu = 0.05 + 0.20 perm / 9,
// sigma = 0.03 + 0.12 perm / 9, weight 0.5 when input and output agree, else 0.25.
extern "C" cn_status cn_generate_example_creature(uint32_t n, float* kernels_fft, double* spatial, float* state,
                                                  float* mu, float* sigma, float* weights, uint32_t* src,
                                                  uint32_t* dst) {
  clear_error();
  if (!is_pow2(n) || n < 4 || n > 8192) return set_error(CN_ERR_UNSUPPORTED, "n = %u must be a power of two in [4, 8192]", n);
  const size_t N = n, P = N * N, h = N / 2;
  for (int k = 0; k < kExK; ++k) {
    const int in = k % kExC, out = k / kExC;
    const float perm = (float)(in * kExC + out);
    if (src) src[k] = (uint32_t)in;
    if (dst) dst[k] = (uint32_t)out;
    if (mu) mu[k] = 0.05f + 0.20f * perm / (float)kExK;
    if (sigma) sigma[k] = 0.03f + 0.12f * perm / (float)kExK;
    if (weights) weights[k] = in == out ? 0.5f : 0.25f;
  }
  if (state) {
    const double half = (double)N / 2.0;
    for (int c = 0; c < kExC; ++c) {
      const double s = kExBlob[c][0], ox = kExBlob[c][1], oy = kExBlob[c][2], inv = 1.0 / (2.0 * s * s);
      for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) {
          const double dx = ((double)j - half) / half - ox, dy = ((double)i - half) / half - oy;
          state[c * P + i * N + j] = (float)std::exp(-(dx * dx + dy * dy) * inv);
        }
    }
  }
  if (!kernels_fft && !spatial) return CN_OK;
  auto build = [&](int k) {
    std::vector<double> re(P), im(P, 0.0);
    double tot = 0.0;
    for (size_t y = 0; y < N; ++y)
      for (size_t x = 0; x < N; ++x) {
        const double v = ex_kernel_value(k, (y + h) & (N - 1), (x + h) & (N - 1), N);
        re[y * N + x] = v;
        tot += v;
      }
    if (tot > 0.0)
      for (size_t i = 0; i < P; ++i) re[i] /= tot;
    if (spatial) std::memcpy(spatial + (size_t)k * P, re.data(), P * sizeof(double));
    if (kernels_fft) {
      fft2d(re, im, N);
      float* o = kernels_fft + (size_t)k * P * 2;
      for (size_t i = 0; i < P; ++i) {
        o[2 * i] = (float)re[i];
        o[2 * i + 1] = (float)im[i];
      }
    }
  };
  std::vector<std::thread> th;
  for (int k = 0; k < kExK; ++k) th.emplace_back(build, k);
  for (auto& t : th) t.join();
  return CN_OK;
}
