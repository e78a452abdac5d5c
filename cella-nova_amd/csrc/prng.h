// Synthetic-creature distribution shared by the host generator (creature.cpp) and the device
// generator (gen_kernels.hip): a counter-based PRNG and the Lenia ring (shell) kernel.
//
// PRNG: splitmix64 finaliser chained over (seed, creature, stream, index); a draw is the top
// 53 bits of the hash as a double in [0, 1).  Every value depends only on its key, so host and
// device produce identical draws in any order.
//
// Ring kernel k with radius R: B ~ U{1,2,3} rings with peaks beta_i ~ U[0.3, 1]; at integer
// offset (dx, dy), r = sqrt(dx^2 + dy^2) / R; for r < 1, s = B r, i = floor(s), f = s - i and
// K = beta_i exp(4 - 1/(f (1 - f))) for 0 < f < 1; K = 0 elsewhere (classic Lenia shell,
// Chan 2018).  The kernel is normalised to sum 1 and centred on the origin with periodic wrap.
#pragma once
#include <stdint.h>
#include <cmath>

#if defined(__HIPCC__)
#define CN_HD __host__ __device__
#else
#define CN_HD
#endif

namespace cn {

enum GenStream : uint64_t { kRings = 1, kBeta = 2, kMu = 3, kSigma = 4, kWeight = 5, kState = 6 };

CN_HD inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

CN_HD inline double uniform01(uint64_t seed, uint64_t creature, uint64_t stream, uint64_t index) {
  const uint64_t h = mix64(mix64(mix64(mix64(seed) + creature) + stream) + index);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
}

struct RingParams {
  double R;
  double beta[3];
  int B;
  int pad;
  double tot;  // sum of the kernel over the grid (filled by the normalisation pass)
};

CN_HD inline RingParams ring_params(uint64_t seed, uint64_t creature, uint32_t k, uint32_t radius) {
  RingParams rp;
  rp.R = (double)radius;
  rp.B = 1 + (int)(3.0 * uniform01(seed, creature, kRings, k));  // U{1,2,3}
  for (int i = 0; i < 3; ++i) rp.beta[i] = 0.3 + 0.7 * uniform01(seed, creature, kBeta, (uint64_t)k * 4 + i);
  rp.pad = 0;
  rp.tot = 0.0;
  return rp;
}

// Unnormalised kernel value at integer offset (dx, dy).
CN_HD inline double ring_value(const RingParams& rp, double dx, double dy) {
  const double r = sqrt(dx * dx + dy * dy) / rp.R;
  if (!(r < 1.0)) return 0.0;
  const double s = rp.B * r;
  const int ri = (int)floor(s);
  const double f = s - ri;
  if (!(f > 0.0 && f < 1.0)) return 0.0;
  return rp.beta[ri] * exp(4.0 - 1.0 / (f * (1.0 - f)));
}

// Seed: U[0,1) on the central side x side patch of every channel (side = min(patch, N)),
// drawn at index (c side + i) side + j of the state stream; zero elsewhere.
struct SeedPatch {
  uint32_t side, origin;
};
inline SeedPatch seed_patch_of(uint32_t N, uint32_t patch) {
  const uint32_t side = patch < N ? patch : N;
  return SeedPatch{side, N / 2 - side / 2};
}
CN_HD inline float seed_value(uint64_t seed, uint64_t creature, uint32_t c, uint32_t i, uint32_t j, uint32_t side) {
  return (float)uniform01(seed, creature, kState, ((uint64_t)c * side + i) * side + j);
}

// Periodic offset of grid index j (origin at 0): j for j < N/2, j - N otherwise.
CN_HD inline double wrap_offset(uint32_t j, uint32_t N) { return j < N / 2 ? (double)j : (double)j - (double)N; }

}  // namespace cn
