// Audio consumer of the rendered frames: per-channel frame statistics and the tonal "sonar ping"
// synthesiser of the reference's video command line (ml-rs/audio.rs), plus its WAV writer.
//
// The synthesiser is sequential host DSP (one sample after another through a low-pass filter
// and a feedback delay line), so it stays on the host; only the frame statistics it reads run
// on the device (cn_creature_render_features, k_render_stats), which replaces the reference's
// host loop over the downloaded RGB24 bytes.  Arithmetic follows audio.rs operation by
// operation in f32 (single rounding per operation, no contraction into fused multiply-adds,
// libm's sinf/expf/powf/tanhf/fmodf), so a frame's samples are those the reference computes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include "common.h"

#pragma clang fp contract(off)

namespace cn {
namespace {

constexpr float kTau = 6.28318530717958647692f;   // std::f32::consts::TAU
constexpr size_t kMaxPings = 10;                   // audio.rs:22
constexpr size_t kPingMinSamples = 250;            // audio.rs:25
constexpr size_t kPingMaxSamples = 3000;           // audio.rs:26
constexpr size_t kCooldownMin = 2;                 // audio.rs:29
constexpr size_t kCooldownMax = 10;                // audio.rs:30
constexpr float kVarChangeThreshold = 0.000001f;   // audio.rs:33
constexpr size_t kDelayBufferSize = 16384;         // audio.rs:36
constexpr size_t kDelayTimeSamples = 6615;         // audio.rs:37
constexpr float kDelayFeedback = 0.35f;            // audio.rs:38
constexpr float kDelayWet = 0.30f;                 // audio.rs:39
constexpr float kLpCutoff = 1400.0f;               // audio.rs:42
constexpr float kTremoloRate = 6.0f;               // audio.rs:45

struct Ping {  // audio.rs:62-78
  float phase, freq_start, freq_end;
  size_t remaining, total_duration;
  float amplitude, pan;
};

// f32 -> integer conversions of Rust's `as`: truncation toward zero, saturating, NaN -> 0.
inline size_t as_usize(float x) {
  if (!(x > 0.0f)) return 0;
  return (size_t)x;
}
inline int16_t as_i16(float x) {
  if (std::isnan(x)) return 0;
  if (x <= -32768.0f) return -32768;
  if (x >= 32767.0f) return 32767;
  return (int16_t)x;
}
// f32::min / f32::max / clamp of Rust (a NaN operand of min/max yields the other one).
inline float fmin_r(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a < b ? a : b)); }
inline float fmax_r(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a > b ? a : b)); }
inline float clamp_r(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

}  // namespace

// Features from exact per-channel byte sums (s1 = sum of bytes, s2 = sum of squared bytes over
// `cells` pixels): the same quantities audio.rs:181-209 accumulates in f64 as sum(b / 255) and
// sum((b / 255)^2), here formed once from the exact integers.
void features_from_sums(const uint64_t s1[3], const uint64_t s2[3], uint64_t cells, float means[3],
                        float variances[3]) {
  const double n = (double)cells;
  for (int ch = 0; ch < 3; ++ch) {
    const double sums = (double)s1[ch] / 255.0;
    const double sum_sqs = (double)s2[ch] / 65025.0;
    means[ch] = (float)(sums / n);
    const float mean_sq = (float)(sum_sqs / n);
    variances[ch] = fmax_r(mean_sq - means[ch] * means[ch], 0.0f);
  }
}

}  // namespace cn

struct cn_audio {  // AudioSynth (audio.rs:104-128)
  uint32_t sample_rate;
  cn::Ping pings[cn::kMaxPings];
  size_t npings;
  size_t cooldown[3];
  float prev_variance[3];
  float base_freq[3];
  bool has_prev_frame;
  float delay_buffer[cn::kDelayBufferSize];
  size_t delay_pos;
  float lp_state_l, lp_state_r;
  float tremolo_phase;
  uint32_t noise_seed;
  float smoothed_activity;

  float rand() {  // audio.rs:168-174
    noise_seed = noise_seed * 1103515245u + 12345u;
    return (float)noise_seed / (float)UINT32_MAX;
  }
  float rand_range(float lo, float hi) { return lo + rand() * (hi - lo); }

  void fire_ping(int ch, float change, const float* means) {  // audio.rs:245-306
    using namespace cn;
    if (npings >= kMaxPings) {
      // Drop the quietest (the first of equal minima, as Iterator::min_by).
      size_t pos = 0;
      for (size_t j = 1; j < npings; ++j)
        if (pings[j].amplitude < pings[pos].amplitude) pos = j;
      for (size_t j = pos + 1; j < npings; ++j) pings[j - 1] = pings[j];
      --npings;
    }
    const float base = base_freq[ch];
    const float brightness = means[ch];
    const float center = base * (0.8f + brightness * 0.5f);
    const bool sweep_up = rand() > 0.5f;
    const float freq_start = sweep_up ? center * 0.6f : center * 1.5f;
    const float freq_end = sweep_up ? center * 1.4f : center * 0.5f;
    const float change_norm = fmin_r(change / 0.00005f, 1.0f);
    const size_t duration = kPingMinSamples + as_usize((float)(kPingMaxSamples - kPingMinSamples) * change_norm);
    const float amplitude = fmax_r(fmin_r(change * 23400.0f, 0.95f), 0.10f);
    const float base_pan = ch == 0 ? 0.15f : (ch == 1 ? 0.5f : 0.85f);
    const float pan = clamp_r(base_pan + rand_range(-0.2f, 0.2f), 0.0f, 1.0f);
    cooldown[ch] = kCooldownMin + as_usize((float)(kCooldownMax - kCooldownMin) * (1.0f - change_norm));
    const float phase = rand_range(0.0f, kTau);
    pings[npings++] = Ping{phase, freq_start, freq_end, duration, duration, amplitude, pan};
  }

  void trigger_pings(const float* means, const float* variances) {  // audio.rs:213-243
    using namespace cn;
    if (!has_prev_frame) {
      for (int ch = 0; ch < 3; ++ch) prev_variance[ch] = variances[ch];
      has_prev_frame = true;
      return;
    }
    float total_change = 0.0f;
    for (int ch = 0; ch < 3; ++ch) {
      if (cooldown[ch] > 0) cooldown[ch] -= 1;
      const float change = std::fabs(variances[ch] - prev_variance[ch]);
      total_change += change;
      prev_variance[ch] = variances[ch];
      if (cooldown[ch] > 0 || change < kVarChangeThreshold) continue;
      fire_ping(ch, change, means);
    }
    smoothed_activity = smoothed_activity * 0.85f + total_change * 0.15f;
  }

  void generate_frame(const float* means, const float* variances, size_t num_samples, int16_t* out) {
    using namespace cn;  // audio.rs:310-396
    trigger_pings(means, variances);
    const float sr = (float)sample_rate;
    const float lp_alpha = 1.0f - std::exp(-kTau * kLpCutoff / sr);
    for (size_t i = 0; i < num_samples; ++i) {
      const float t = (float)i;
      const float tremolo = std::sin(kTau * kTremoloRate * t / sr + tremolo_phase) * 0.15f + 0.85f;
      float left = 0.0f, right = 0.0f;
      for (size_t j = 0; j < npings; ++j) {
        Ping& p = pings[j];
        if (p.remaining == 0) continue;
        const float progress = (float)p.remaining / (float)p.total_duration;
        float env = progress > 0.93f ? (1.0f - progress) * 14.0f : std::pow(progress / 0.93f, 1.6f);
        env = clamp_r(env, 0.0f, 1.0f);
        const float tf = 1.0f - ((float)p.remaining / (float)p.total_duration);
        const float freq = p.freq_start + (p.freq_end - p.freq_start) * tf;
        const float phase_inc = kTau * freq / sr;
        const float tone = std::sin(p.phase);
        p.phase += phase_inc;
        if (p.phase >= kTau) p.phase -= kTau;
        const float harm2 = std::sin(p.phase * 2.0f) * 0.15f * env;
        const float sample = (tone + harm2) * env * p.amplitude * tremolo;
        left += sample * (1.0f - p.pan);
        right += sample * p.pan;
        p.remaining = p.remaining > 0 ? p.remaining - 1 : 0;
      }
      size_t keep = 0;  // retain the live pings, in order
      for (size_t j = 0; j < npings; ++j)
        if (pings[j].remaining > 0) pings[keep++] = pings[j];
      npings = keep;

      lp_state_l += lp_alpha * (left - lp_state_l);
      lp_state_r += lp_alpha * (right - lp_state_r);
      const float filtered_l = lp_state_l, filtered_r = lp_state_r;

      const size_t delay_read = (delay_pos + kDelayBufferSize - kDelayTimeSamples) % kDelayBufferSize;
      const float echo_l = delay_buffer[delay_read];
      const float echo_r = delay_buffer[(delay_read + 1) % kDelayBufferSize];
      const float wet_l = filtered_l + echo_l * kDelayWet;
      const float wet_r = filtered_r + echo_r * kDelayWet;
      delay_buffer[delay_pos] = filtered_l + echo_l * kDelayFeedback;
      delay_buffer[(delay_pos + 1) % kDelayBufferSize] = filtered_r + echo_r * kDelayFeedback;
      delay_pos = (delay_pos + 2) % kDelayBufferSize;

      const float out_l = std::tanh(wet_l * 5.0f);
      const float out_r = std::tanh(wet_r * 5.0f);
      out[2 * i] = as_i16(clamp_r(out_l * 32767.0f, -32768.0f, 32767.0f));
      out[2 * i + 1] = as_i16(clamp_r(out_r * 32767.0f, -32768.0f, 32767.0f));
    }
    tremolo_phase += kTau * kTremoloRate * (float)num_samples / sr;
    tremolo_phase = std::fmod(tremolo_phase, kTau);
  }
};

using namespace cn;

extern "C" {

cn_status cn_audio_extract_features(const uint8_t* rgb, uint32_t grid, float* means, float* variances) {
  CN_TRY
  clear_error();
  if (!rgb || !means || !variances) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (grid == 0) return set_error(CN_ERR_INVALID_ARG, "grid size 0");
  const uint64_t cells = (uint64_t)grid * grid;
  uint64_t s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
  for (uint64_t i = 0; i < cells; ++i)
    for (int ch = 0; ch < 3; ++ch) {
      const uint64_t b = rgb[3 * i + ch];
      s1[ch] += b;
      s2[ch] += b * b;
    }
  features_from_sums(s1, s2, cells, means, variances);
  return CN_OK;
  CN_CATCH
}

cn_status cn_audio_create(uint32_t sample_rate, cn_audio** out) {
  CN_TRY
  clear_error();
  if (!out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (sample_rate == 0) return set_error(CN_ERR_INVALID_ARG, "sample rate 0");
  cn_audio* a = new (std::nothrow) cn_audio();
  if (!a) return set_error(CN_ERR_OUT_OF_MEMORY, "audio synthesiser");
  a->sample_rate = sample_rate;  // AudioSynth::new (audio.rs:131-160)
  a->npings = 0;
  for (int ch = 0; ch < 3; ++ch) {
    a->cooldown[ch] = 0;
    a->prev_variance[ch] = 0.0f;
  }
  a->base_freq[0] = 100.0f;
  a->base_freq[1] = 300.0f;
  a->base_freq[2] = 750.0f;
  a->has_prev_frame = false;
  std::memset(a->delay_buffer, 0, sizeof(a->delay_buffer));
  a->delay_pos = 0;
  a->lp_state_l = a->lp_state_r = 0.0f;
  a->tremolo_phase = 0.0f;
  a->noise_seed = 0xC0FFEEu;
  a->smoothed_activity = 0.0f;
  *out = a;
  return CN_OK;
  CN_CATCH
}

cn_status cn_audio_destroy(cn_audio* a) {
  delete a;
  return CN_OK;
}

cn_status cn_audio_generate_frame(cn_audio* a, const float* means, const float* variances, uint32_t num_samples,
                                  int16_t* out, size_t n) {
  CN_TRY
  clear_error();
  if (!a || !means || !variances || (!out && num_samples)) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (n != 2 * (size_t)num_samples)
    return set_error(CN_ERR_INVALID_ARG, "output holds %zu samples, expected %zu (stereo)", n, 2 * (size_t)num_samples);
  a->generate_frame(means, variances, num_samples, out);
  return CN_OK;
  CN_CATCH
}

cn_status cn_audio_activity(const cn_audio* a, float* out) {
  CN_TRY
  clear_error();
  if (!a || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  *out = a->smoothed_activity;
  return CN_OK;
  CN_CATCH
}

cn_status cn_write_wav(const char* path, const int16_t* samples, size_t count, uint32_t sample_rate,
                       uint16_t channels) {
  CN_TRY
  clear_error();
  if (!path || (!samples && count)) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (channels == 0 || sample_rate == 0) return set_error(CN_ERR_INVALID_ARG, "channels and sample rate must be positive");
  if (count > (0xFFFFFFFFull - 36) / 2) return set_error(CN_ERR_INVALID_ARG, "%zu samples exceed a WAV file", count);
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(CN_ERR_IO, "cannot create %s", path);
  // RIFF/WAVE PCM header (audio.rs:412-441), little-endian fields.
  uint8_t h[44];
  auto le32 = [&](int at, uint32_t v) {
    for (int j = 0; j < 4; ++j) h[at + j] = (uint8_t)(v >> (8 * j));
  };
  auto le16 = [&](int at, uint16_t v) {
    h[at] = (uint8_t)v;
    h[at + 1] = (uint8_t)(v >> 8);
  };
  const uint32_t data_size = (uint32_t)(count * 2);
  std::memcpy(h, "RIFF", 4);
  le32(4, 36 + data_size);
  std::memcpy(h + 8, "WAVEfmt ", 8);
  le32(16, 16);
  le16(20, 1);
  le16(22, channels);
  le32(24, sample_rate);
  le32(28, sample_rate * channels * 2);
  le16(32, (uint16_t)(channels * 2));
  le16(34, 16);
  std::memcpy(h + 36, "data", 4);
  le32(40, data_size);
  bool ok = std::fwrite(h, 1, 44, f) == 44;
  uint8_t buf[4096];
  for (size_t i = 0; ok && i < count;) {
    size_t m = 0;
    for (; m + 2 <= sizeof(buf) && i < count; ++i, m += 2) {
      const uint16_t v = (uint16_t)samples[i];
      buf[m] = (uint8_t)v;
      buf[m + 1] = (uint8_t)(v >> 8);
    }
    ok = std::fwrite(buf, 1, m, f) == m;
  }
  if (std::fclose(f) != 0) ok = false;
  if (!ok) return set_error(CN_ERR_IO, "write failed: %s", path);
  return CN_OK;
  CN_CATCH
}

}  // extern "C"
