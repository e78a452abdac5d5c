// Host helpers shared by the single-creature engine and the batch runner.
#pragma once
#include <vector>
#include "common.h"
#include "types.h"

namespace cn {

// kTwN-entry twiddle table exp(-2*pi*i*j/kTwN), interleaved (re, im), FP64-computed.
std::vector<float> make_twiddles();

// Host-side validation of one creature's parameters (before any HIP call).
cn_status validate_creature(uint32_t n, uint32_t C, uint32_t K, const uint32_t* src, const uint32_t* dst,
                            const float* mu, const float* sigma, const float* w, float temp);

cn_status validate_growth(uint32_t K, const float* mu, const float* sigma, const float* w);
void fill_growth(CreatureDev& d, uint32_t K, const float* mu, const float* sigma, const float* w);

// Fills the CSR kernel lists and growth parameters of a descriptor.
void fill_desc_params(CreatureDev& d, uint32_t C, uint32_t K, const uint32_t* src, const uint32_t* dst,
                      const float* mu, const float* sigma, const float* w, float temp);

// Full N*N spectrum (interleaved) -> Hermitian part / N^2 in the half layout.
void project_kernel(const float* full, uint32_t N, float* half, float* nyq);

}  // namespace cn
