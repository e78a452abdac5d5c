// Host-side helpers shared by the C ABI translation units: thread-local error text, the
// checked HIP-call wrapper with a fault-injection counter, and small validation utilities.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <string>
#include "../../include/cellanova.h"

namespace cn {

// Error text of the calling thread (errors.cpp): a fixed buffer of kErrorTextSize bytes, so
// recording an error never allocates (it also works while reporting a failed allocation).
constexpr size_t kErrorTextSize = 1024;
cn_status set_error(cn_status st, const char* fmt, ...);
void clear_error();
const char* last_error_text();
// Inside a catch block: records the exception in flight as a status (std::bad_alloc and
// std::length_error as CN_ERR_OUT_OF_MEMORY) and returns it.
cn_status status_of_exception() noexcept;

// Wraps the body of every C ABI function that returns cn_status: no C++ exception crosses the
// boundary.
#define CN_TRY try {
#define CN_CATCH                          \
  }                                       \
  catch (...) {                           \
    return ::cn::status_of_exception();   \
  }

// Returns true (and records CN_ERR_HIP) if the call failed or a failure was injected.
bool hip_failed(hipError_t err, const char* what, cn_status* st);
bool inject_failure();  // consumes one injected failure if any is pending

#define CN_HIP(expr)                                                                  \
  do {                                                                                \
    cn_status _st;                                                                    \
    if (::cn::inject_failure()) return ::cn::set_error(CN_ERR_HIP, "%s: injected failure", #expr); \
    if (::cn::hip_failed((expr), #expr, &_st)) return _st;                            \
  } while (0)

inline bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
inline bool finite_all(const float* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

// Synthetic creatures (creature.cpp): parameter checks and growth parameters of the
// benchmark distribution, shared by the host and device generators.
cn_status check_creature_params(const cn_creature_params* p);

// Audio features (means and variances of bytes / 255 per RGB channel) from exact byte sums over
// `cells` pixels (audio.cpp).
void features_from_sums(const uint64_t s1[3], const uint64_t s2[3], uint64_t cells, float means[3],
                        float variances[3]);
void generate_growth(const cn_creature_params* p, float* mu, float* sigma, float* weights);

}  // namespace cn
