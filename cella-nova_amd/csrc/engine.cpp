// Engine: the MI355X counterpart of the reference's GpuMaceLenia (ml-rs/orchestrator/mod.rs).
// It owns the device buffers of one creature, builds the device descriptor, runs the step
// (four kernels, captured once into two HIP graphs for the two ping-pong parities) and
// implements the C ABI entry points of include/cellanova.h that act on an engine.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "common.h"
#include "engine.h"
#include "launch.h"

namespace cn {

// ---------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------
static thread_local std::string g_err;
static int g_fail_next = 0;

cn_status set_error(cn_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}
void clear_error() { g_err.clear(); }

bool inject_failure() {
  if (g_fail_next > 0) {
    --g_fail_next;
    return true;
  }
  return false;
}

bool hip_failed(hipError_t err, const char* what, cn_status* st) {
  if (err == hipSuccess) return false;
  const cn_status s = (err == hipErrorOutOfMemory) ? CN_ERR_OUT_OF_MEMORY : CN_ERR_HIP;
  *st = set_error(s, "%s failed: %s (%d)", what, hipGetErrorString(err), (int)err);
  return true;
}

// ---------------------------------------------------------------------------------------
// twiddles: tw[i] = exp(-2*pi*i*i/kTwN), FP64 then rounded once
// ---------------------------------------------------------------------------------------
std::vector<float> make_twiddles() {
  std::vector<float> t(2 * kTwN);
  for (int i = 0; i < kTwN; ++i) {
    const double a = -2.0 * M_PI * (double)i / (double)kTwN;
    t[2 * i] = (float)std::cos(a);
    t[2 * i + 1] = (float)std::sin(a);
  }
  return t;
}

cn_status validate_creature(uint32_t n, uint32_t C, uint32_t K, const uint32_t* src, const uint32_t* dst,
                            const float* mu, const float* sigma, const float* w, float temp) {
  if (!is_pow2(n) || n < 16 || n > 2048)
    return set_error(CN_ERR_UNSUPPORTED, "grid size %u must be a power of two in [16, 2048]", n);
  if (C < 1 || C > (uint32_t)kMaxChannels)
    return set_error(CN_ERR_UNSUPPORTED, "num_channels %u must be in [1, %d]", C, kMaxChannels);
  if (K < 1 || K > (uint32_t)kMaxKernels)
    return set_error(CN_ERR_UNSUPPORTED, "num_kernels %u must be in [1, %d]", K, kMaxKernels);
  if (!src || !dst || !mu || !sigma || !w) return set_error(CN_ERR_INVALID_ARG, "null parameter array");
  for (uint32_t k = 0; k < K; ++k) {
    if (src[k] >= C) return set_error(CN_ERR_INVALID_ARG, "src[%u] = %u out of range (C = %u)", k, src[k], C);
    if (dst[k] >= C) return set_error(CN_ERR_INVALID_ARG, "dst[%u] = %u out of range (C = %u)", k, dst[k], C);
    if (!std::isfinite(sigma[k]) || !(sigma[k] > 0.f))
      return set_error(CN_ERR_INVALID_ARG, "sigma[%u] = %g must be finite and > 0", k, (double)sigma[k]);
    if (!std::isfinite(mu[k])) return set_error(CN_ERR_INVALID_ARG, "mu[%u] is not finite", k);
    if (!std::isfinite(w[k])) return set_error(CN_ERR_INVALID_ARG, "weights[%u] is not finite", k);
  }
  if (!std::isfinite(temp)) return set_error(CN_ERR_INVALID_ARG, "temp is not finite");
  return CN_OK;
}

void fill_desc_params(CreatureDev& d, uint32_t C, uint32_t K, const uint32_t* src, const uint32_t* dst,
                      const float* mu, const float* sigma, const float* w, float temp) {
  d.C = (int)C;
  d.K = (int)K;
  d.temp = temp;
  int pos = 0;
  for (uint32_t c = 0; c < C; ++c) {
    d.src_off[c] = pos;
    for (uint32_t k = 0; k < K; ++k)
      if (src[k] == c) d.src_k[pos++] = (int)k;
  }
  for (uint32_t c = C; c <= (uint32_t)kMaxChannels; ++c) d.src_off[c] = pos;
  pos = 0;
  for (uint32_t c = 0; c < C; ++c) {
    d.dst_off[c] = pos;
    for (uint32_t k = 0; k < K; ++k)
      if (dst[k] == c) d.dst_k[pos++] = (int)k;
  }
  for (uint32_t c = C; c <= (uint32_t)kMaxChannels; ++c) d.dst_off[c] = pos;
  for (uint32_t k = 0; k < K; ++k) {
    d.mu[k] = mu[k];
    d.inv_s2[k] = 1.0f / (sigma[k] * sigma[k]);  // example.rs:247
    d.w[k] = w[k];
  }
}

// Hermitian part of a full N*N spectrum, scaled by 1/N^2, split into the half layout:
// half[y][x] for x < N/2 and nyq[y] for x = N/2.
void project_kernel(const float* full, uint32_t N, float* half, float* nyq) {
  const double scale = 1.0 / ((double)N * (double)N);
  const uint32_t H = N / 2, m = N - 1;
  auto herm = [&](uint32_t y, uint32_t x, float* o) {
    const size_t a = ((size_t)y * N + x) * 2, b = ((size_t)((N - y) & m) * N + ((N - x) & m)) * 2;
    const double re = 0.5 * ((double)full[a] + (double)full[b]);
    const double im = 0.5 * ((double)full[a + 1] - (double)full[b + 1]);
    o[0] = (float)(re * scale);
    o[1] = (float)(im * scale);
  };
  for (uint32_t y = 0; y < N; ++y) {
    for (uint32_t x = 0; x < H; ++x) herm(y, x, half + ((size_t)y * H + x) * 2);
    herm(y, H, nyq + (size_t)y * 2);
  }
}

}  // namespace cn

using namespace cn;

struct cn_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t N = 0, C = 0, K = 0;
  int cur = 0;  // state[cur] holds the current world
  uint64_t steps = 0;
  bool have_prev = false;  // state[cur ^ 1] holds the input of the most recent step
  uint64_t bytes = 0;
  CreatureDev host_desc{};
  CreatureDev* d_desc = nullptr;
  float2* tw = nullptr;
  double2* mass_partial = nullptr;
  int mass_nblk = 0;
  std::vector<void*> allocs;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int col_tc = 0;  // fused column pass tile width (0: built-in default for N)
};

namespace {

cn_status dalloc(cn_engine* e, void** p, size_t bytes) {
  CN_HIP(hipMalloc(p, bytes));
  e->allocs.push_back(*p);
  e->bytes += bytes;
  CN_HIP(hipMemsetAsync(*p, 0, bytes, e->stream));
  return CN_OK;
}

void release(cn_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  for (auto& g : e->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& v : e->ev)
    if (v) (void)hipEventDestroy(v);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (void* p : e->allocs) (void)hipFree(p);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

StepLaunch base_launch(cn_engine* e, int parity) {
  StepLaunch sl;
  sl.descs = e->d_desc;
  sl.count = 1;
  sl.parity = parity;
  sl.max_c = (int)e->C;
  sl.tw = e->tw;
  sl.col_tc = e->col_tc;
  return sl;
}

cn_status enqueue_step(cn_engine* e) {
  if (!e->graph[e->cur]) {
    hipGraph_t g = nullptr;
    CN_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t le = launch_step((int)e->N, base_launch(e, e->cur), e->stream);
    const hipError_t ce = hipStreamEndCapture(e->stream, &g);
    cn_status st;
    if (hip_failed(le, "launch_step (capture)", &st)) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (hip_failed(ce, "hipStreamEndCapture", &st)) return st;
    const hipError_t ie = hipGraphInstantiate(&e->graph[e->cur], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (hip_failed(ie, "hipGraphInstantiate", &st)) return st;
  }
  CN_HIP(hipGraphLaunch(e->graph[e->cur], e->stream));
  e->cur ^= 1;
  e->steps++;
  e->have_prev = true;
  return CN_OK;
}

}  // namespace

extern "C" {

const char* cn_last_error(void) { return g_err.c_str(); }

void cn_debug_fail_next_hip(int count) { g_fail_next = count > 0 ? count : 0; }

cn_status cn_engine_create(const cn_engine_desc* desc, cn_engine** out) {
  clear_error();
  if (!desc || !out) return set_error(CN_ERR_INVALID_ARG, "null desc or out");
  *out = nullptr;
  cn_status st = validate_creature(desc->n, desc->num_channels, desc->num_kernels, desc->src, desc->dst,
                                   desc->mu, desc->sigma, desc->weights, desc->temp);
  if (st != CN_OK) return st;
  if (desc->device < 0) return set_error(CN_ERR_INVALID_ARG, "device %d", desc->device);

  cn_engine* e = new cn_engine();
  e->device = desc->device;
  e->N = desc->n;
  if (const char* tc = std::getenv("CN_COL_TC")) e->col_tc = std::atoi(tc);
  e->C = desc->num_channels;
  e->K = desc->num_kernels;
  auto fail = [&](cn_status s) {
    release(e);
    return s;
  };
  {
    cn_status s2;
    if (inject_failure()) return fail(set_error(CN_ERR_HIP, "hipSetDevice: injected failure"));
    if (hip_failed(hipSetDevice(e->device), "hipSetDevice", &s2)) return fail(s2);
    if (hip_failed(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "hipStreamCreate", &s2))
      return fail(s2);
    for (auto& v : e->ev)
      if (hip_failed(hipEventCreate(&v), "hipEventCreate", &s2)) return fail(s2);
  }
  const size_t N = e->N, C = e->C, K = e->K, P = N * N, H = N / 2;
  CreatureDev& d = e->host_desc;
  d.N = (int)N;
  void* p = nullptr;
  for (int s = 0; s < 2; ++s) {
    if ((st = dalloc(e, &p, C * P * 4)) != CN_OK) return fail(st);
    d.state[s] = (float*)p;
  }
  if ((st = dalloc(e, &p, C * N * H * 8)) != CN_OK) return fail(st);
  d.spec = p;
  if ((st = dalloc(e, &p, K * N * H * 8)) != CN_OK) return fail(st);
  d.khalf = p;
  if ((st = dalloc(e, &p, K * N * 8)) != CN_OK) return fail(st);
  d.knyq = p;
  if ((st = dalloc(e, &p, K * N * H * 8)) != CN_OK) return fail(st);
  d.uhat = p;
  if ((st = dalloc(e, &p, C * P * 4)) != CN_OK) return fail(st);
  d.A = (float*)p;
  d.u = nullptr;  // allocated on first conv readback
  fill_desc_params(d, e->C, e->K, desc->src, desc->dst, desc->mu, desc->sigma, desc->weights, desc->temp);
  if ((st = dalloc(e, &p, sizeof(CreatureDev))) != CN_OK) return fail(st);
  e->d_desc = (CreatureDev*)p;
  {
    cn_status s2;
    if (inject_failure()) return fail(set_error(CN_ERR_HIP, "hipMemcpy: injected failure"));
    if (hip_failed(hipMemcpyAsync(e->d_desc, &d, sizeof d, hipMemcpyHostToDevice, e->stream), "hipMemcpy desc",
                   &s2))
      return fail(s2);
  }
  if ((st = dalloc(e, &p, (size_t)kTwN * 8)) != CN_OK) return fail(st);
  e->tw = (float2*)p;
  const std::vector<float> tw = make_twiddles();
  e->mass_nblk = 64;
  if ((st = dalloc(e, &p, (size_t)kMaxChannels * e->mass_nblk * sizeof(double2))) != CN_OK) return fail(st);
  e->mass_partial = (double2*)p;
  {
    cn_status s2;
    if (hip_failed(hipMemcpyAsync(e->tw, tw.data(), tw.size() * 4, hipMemcpyHostToDevice, e->stream),
                   "hipMemcpy twiddles", &s2))
      return fail(s2);
    if (hip_failed(hipStreamSynchronize(e->stream), "hipStreamSynchronize", &s2)) return fail(s2);
  }
  *out = e;
  return CN_OK;
}

cn_status cn_engine_destroy(cn_engine* e) {
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  release(e);
  return CN_OK;
}

cn_status cn_set_kernel(cn_engine* e, uint32_t k, const float* interleaved, size_t n_complex) {
  clear_error();
  if (!e || !interleaved) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (k >= e->K) return set_error(CN_ERR_INVALID_ARG, "kernel index %u out of range (K = %u)", k, e->K);
  const size_t N = e->N, H = N / 2;
  if (n_complex != N * N)
    return set_error(CN_ERR_INVALID_ARG, "kernel spectrum has %zu values, expected %zu", n_complex, N * N);
  if (!finite_all(interleaved, 2 * n_complex)) return set_error(CN_ERR_INVALID_ARG, "kernel spectrum not finite");
  std::vector<float> half(N * H * 2), nyq(N * 2);
  project_kernel(interleaved, e->N, half.data(), nyq.data());
  CN_HIP(hipSetDevice(e->device));
  const CreatureDev& d = e->host_desc;
  CN_HIP(hipMemcpyAsync((char*)d.khalf + (size_t)k * N * H * 8, half.data(), half.size() * 4,
                        hipMemcpyHostToDevice, e->stream));
  CN_HIP(hipMemcpyAsync((char*)d.knyq + (size_t)k * N * 8, nyq.data(), nyq.size() * 4, hipMemcpyHostToDevice,
                        e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  return CN_OK;
}

static cn_status upload_f32(cn_engine* e, uint32_t c, const float* data) {
  CN_HIP(hipSetDevice(e->device));
  const size_t P = (size_t)e->N * e->N;
  CN_HIP(hipMemcpyAsync(e->host_desc.state[e->cur] + c * P, data, P * 4, hipMemcpyHostToDevice, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  e->have_prev = false;
  return CN_OK;
}

cn_status cn_upload_channel_f64(cn_engine* e, uint32_t channel, const double* data, size_t n) {
  clear_error();
  if (!e || !data) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (channel >= e->C) return set_error(CN_ERR_INVALID_ARG, "channel %u out of range (C = %u)", channel, e->C);
  const size_t P = (size_t)e->N * e->N;
  if (n != P) return set_error(CN_ERR_INVALID_ARG, "channel has %zu values, expected %zu", n, P);
  std::vector<float> f(P);
  for (size_t i = 0; i < P; ++i) {
    if (!std::isfinite(data[i])) return set_error(CN_ERR_INVALID_ARG, "channel value %zu not finite", i);
    f[i] = (float)data[i];  // mod.rs:220
  }
  return upload_f32(e, channel, f.data());
}

cn_status cn_upload_channel_f32(cn_engine* e, uint32_t channel, const float* data, size_t n) {
  clear_error();
  if (!e || !data) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (channel >= e->C) return set_error(CN_ERR_INVALID_ARG, "channel %u out of range (C = %u)", channel, e->C);
  const size_t P = (size_t)e->N * e->N;
  if (n != P) return set_error(CN_ERR_INVALID_ARG, "channel has %zu values, expected %zu", n, P);
  if (!finite_all(data, P)) return set_error(CN_ERR_INVALID_ARG, "channel values not finite");
  return upload_f32(e, channel, data);
}

cn_status cn_iterate(cn_engine* e) {
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  CN_HIP(hipSetDevice(e->device));
  return enqueue_step(e);
}

cn_status cn_run_steps(cn_engine* e, uint64_t n) {
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  CN_HIP(hipSetDevice(e->device));
  for (uint64_t i = 0; i < n; ++i) {
    const cn_status st = enqueue_step(e);
    if (st != CN_OK) return st;
  }
  return CN_OK;
}

cn_status cn_sync(cn_engine* e) {
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipStreamSynchronize(e->stream));
  return CN_OK;
}

// One directly launched step with events around every kernel; kernel_ms receives
// [row forward, fused column, inverse row + growth, diffusion] in milliseconds.
static cn_status timed_step(cn_engine* e, float kernel_ms[4]) {
  StepLaunch sl = base_launch(e, e->cur);
  for (int i = 0; i < 5; ++i) sl.ev[i] = e->ev[i];
  CN_HIP(launch_step((int)e->N, sl, e->stream));
  CN_HIP(hipEventSynchronize(e->ev[4]));
  for (int i = 0; i < 4; ++i) CN_HIP(hipEventElapsedTime(&kernel_ms[i], e->ev[i], e->ev[i + 1]));
  e->cur ^= 1;
  e->steps++;
  e->have_prev = true;
  return CN_OK;
}

cn_status cn_timed_iterate(cn_engine* e, double ms[4]) {
  clear_error();
  if (!e || !ms) return set_error(CN_ERR_INVALID_ARG, "null argument");
  CN_HIP(hipSetDevice(e->device));
  float k[4];
  const cn_status st = timed_step(e, k);
  if (st != CN_OK) return st;
  ms[0] = (double)k[0] + (double)k[1];  // convolution: row pass + fused column pass
  ms[1] = k[2];                         // inverse row pass fused with growth and affinity
  ms[2] = k[3];                         // diffusion
  ms[3] = 0.0;                          // copy: ping-pong pointer swap
  return CN_OK;
}

cn_status cn_profile_kernels(cn_engine* e, uint64_t steps, double* kernel_ms, size_t n) {
  clear_error();
  if (!e || !kernel_ms) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (n != 4) return set_error(CN_ERR_INVALID_ARG, "kernel_ms must have 4 entries");
  if (steps == 0) return set_error(CN_ERR_INVALID_ARG, "steps must be > 0");
  CN_HIP(hipSetDevice(e->device));
  double acc[4] = {0, 0, 0, 0};
  for (uint64_t s = 0; s < steps; ++s) {
    float k[4];
    const cn_status st = timed_step(e, k);
    if (st != CN_OK) return st;
    for (int i = 0; i < 4; ++i) acc[i] += k[i];
  }
  for (int i = 0; i < 4; ++i) kernel_ms[i] = acc[i] / (double)steps;
  return CN_OK;
}

cn_status cn_download_channels(cn_engine* e, float* out, size_t n) {
  clear_error();
  if (!e || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  const size_t tot = (size_t)e->C * e->N * e->N;
  if (n != tot) return set_error(CN_ERR_INVALID_ARG, "output has %zu values, expected %zu", n, tot);
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipMemcpyAsync(out, e->host_desc.state[e->cur], tot * 4, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  return CN_OK;
}

cn_status cn_download_conv_real(cn_engine* e, float* out, size_t n) {
  clear_error();
  if (!e || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  const size_t P = (size_t)e->N * e->N, tot = (size_t)e->K * P;
  if (n != tot) return set_error(CN_ERR_INVALID_ARG, "output has %zu values, expected %zu", n, tot);
  if (!e->have_prev) {
    if (e->steps == 0) {
      std::memset(out, 0, tot * 4);
      return CN_OK;
    }
    return set_error(CN_ERR_STATE, "the input of the last step was overwritten by an upload");
  }
  CN_HIP(hipSetDevice(e->device));
  CreatureDev& d = e->host_desc;
  if (!d.u) {
    void* p = nullptr;
    const cn_status st = dalloc(e, &p, tot * 4);
    if (st != CN_OK) return st;
    d.u = (float*)p;
    CN_HIP(hipMemcpyAsync(e->d_desc, &d, sizeof d, hipMemcpyHostToDevice, e->stream));
  }
  // Recompute the convolution of the state the most recent step started from.
  StepLaunch sl = base_launch(e, e->cur ^ 1);
  sl.write_u = true;
  sl.skip_diffusion = true;
  CN_HIP(launch_step((int)e->N, sl, e->stream));
  CN_HIP(hipMemcpyAsync(out, d.u, tot * 4, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  const float scale = (float)P;  // unnormalised like the reference's conv buffer (exact: power of 2)
  for (size_t i = 0; i < tot; ++i) out[i] *= scale;
  return CN_OK;
}

cn_status cn_mass_fp64(cn_engine* e, double* out, size_t n_channels) {
  clear_error();
  if (!e || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (n_channels != e->C) return set_error(CN_ERR_INVALID_ARG, "expected %u channels", e->C);
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(launch_mass(e->d_desc, 1, e->cur, (int)e->C, e->mass_partial, e->mass_nblk, e->stream));
  std::vector<double> part((size_t)kMaxChannels * e->mass_nblk * 2);
  CN_HIP(hipMemcpyAsync(part.data(), e->mass_partial, part.size() * 8, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  for (uint32_t c = 0; c < e->C; ++c) {
    double s = 0.0, comp = 0.0;
    for (int b = 0; b < e->mass_nblk; ++b) {
      const double* pr = &part[((size_t)c * e->mass_nblk + b) * 2];
      for (int j = 0; j < 2; ++j) {
        const double x = pr[j], t = s + x;
        comp += (std::fabs(s) >= std::fabs(x)) ? (s - t) + x : (x - t) + s;
        s = t;
      }
    }
    out[c] = s + comp;
  }
  return CN_OK;
}

cn_status cn_engine_get_info(const cn_engine* e, cn_engine_info* out) {
  clear_error();
  if (!e || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  out->n = e->N;
  out->num_channels = e->C;
  out->num_kernels = e->K;
  out->steps_done = e->steps;
  out->device_bytes = e->bytes;
  return CN_OK;
}

}  // extern "C"
