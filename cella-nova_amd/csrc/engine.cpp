// Engine: the MI355X counterpart of the reference's GpuMaceLenia (ml-rs/orchestrator/mod.rs),
// generalised to a batch of creatures.  One engine owns the device buffers of one or more
// creatures (each with its own N, C, K, maps and parameters), grouped into size classes (one
// descriptor array and one launch per phase per class), one HIP stream, and HIP graphs:
//   - the step graph (all classes, one per ping-pong parity) replayed by run_steps;
//   - the rollout graph (forward T steps with checkpoints, MSE loss, checkpointed adjoint)
//     captured once per (T, checkpoint stride) and replayed per optimisation iteration.
// The single-creature C ABI of include/cellanova.h is the batch API with creature 0.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <type_traits>
#include <vector>
#include "common.h"
#include "engine.h"
#include "cfg.h"
#include "launch.h"
#include "mass.h"

namespace cn {

// ---------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------
static int g_fail_next = 0;
static int g_fail_at = -1;  // checked calls to let through before one injected failure

bool inject_failure() {
  if (g_fail_next > 0) {
    --g_fail_next;
    return true;
  }
  if (g_fail_at >= 0) return g_fail_at-- == 0;
  return false;
}

bool hip_failed(hipError_t err, const char* what, cn_status* st) {
  if (err == hipSuccess) return false;
  const cn_status s = (err == hipErrorOutOfMemory) ? CN_ERR_OUT_OF_MEMORY : CN_ERR_HIP;
  *st = set_error(s, "%s failed: %s (%d)", what, hipGetErrorString(err), (int)err);
  return true;
}

// ---------------------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------------------
// Twiddle table (layout in types.h): the kTwN base entries computed in FP64 and rounded once,
// then the per-pass blocks as copies of base entries.
std::vector<float> make_twiddles() {
  std::vector<float> t(2 * (size_t)tw_table_size());
  for (int i = 0; i < kTwN; ++i) {
    const double a = -2.0 * M_PI * (double)i / (double)kTwN;
    t[2 * i] = (float)std::cos(a);
    t[2 * i + 1] = (float)std::sin(a);
  }
  for (int r = 2; r <= 16; r *= 2)
    for (int ns = 2; ns * r <= kTwN; ns *= 2) {
      const int off = tw_pass_offset(ns, r), stride = kTwN / (ns * r);
      for (int m = 1; m < r; ++m)
        for (int k = 0; k < ns; ++k) {
          const int src = (m * k * stride) % kTwN, dst = off + (m - 1) * ns + k;
          t[2 * dst] = t[2 * src];
          t[2 * dst + 1] = t[2 * src + 1];
        }
    }
  return t;
}

cn_status validate_growth(uint32_t K, const float* mu, const float* sigma, const float* w) {
  if (!mu || !sigma || !w) return set_error(CN_ERR_INVALID_ARG, "null growth parameter array");
  for (uint32_t k = 0; k < K; ++k) {
    if (!std::isfinite(sigma[k]) || !(sigma[k] > 0.f))
      return set_error(CN_ERR_INVALID_ARG, "sigma[%u] = %g must be finite and > 0", k, (double)sigma[k]);
    if (!std::isfinite(mu[k])) return set_error(CN_ERR_INVALID_ARG, "mu[%u] is not finite", k);
    if (!std::isfinite(w[k])) return set_error(CN_ERR_INVALID_ARG, "weights[%u] is not finite", k);
  }
  return CN_OK;
}

cn_status validate_creature(uint32_t n, uint32_t C, uint32_t K, const uint32_t* src, const uint32_t* dst,
                            const float* mu, const float* sigma, const float* w, float temp) {
  if (!is_pow2(n) || n < 16 || n > 2048)
    return set_error(CN_ERR_UNSUPPORTED, "grid size %u must be a power of two in [16, 2048]", n);
  if (C < 1 || C > (uint32_t)kMaxChannels)
    return set_error(CN_ERR_UNSUPPORTED, "num_channels %u must be in [1, %d]", C, kMaxChannels);
  if (K < 1 || K > (uint32_t)kMaxKernels)
    return set_error(CN_ERR_UNSUPPORTED, "num_kernels %u must be in [1, %d]", K, kMaxKernels);
  if (!src || !dst) return set_error(CN_ERR_INVALID_ARG, "null src/dst array");
  for (uint32_t k = 0; k < K; ++k) {
    if (src[k] >= C) return set_error(CN_ERR_INVALID_ARG, "src[%u] = %u out of range (C = %u)", k, src[k], C);
    if (dst[k] >= C) return set_error(CN_ERR_INVALID_ARG, "dst[%u] = %u out of range (C = %u)", k, dst[k], C);
  }
  const cn_status st = validate_growth(K, mu, sigma, w);
  if (st != CN_OK) return st;
  if (!std::isfinite(temp)) return set_error(CN_ERR_INVALID_ARG, "temp is not finite");
  return CN_OK;
}

void fill_growth(CreatureDev& d, uint32_t K, const float* mu, const float* sigma, const float* w) {
  for (uint32_t k = 0; k < K; ++k) {
    d.mu[k] = mu[k];
    d.sigma[k] = sigma[k];
    d.inv_s2[k] = 1.0f / (sigma[k] * sigma[k]);  // example.rs:247
    d.w[k] = w[k];
  }
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
    d.src_of[k] = (int)src[k];
    d.dst_of[k] = (int)dst[k];
  }
  fill_growth(d, K, mu, sigma, w);
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

// True if the (Hermitian-projected) spectrum is real and even in ky to within f32 rounding
// of its largest coefficient (|Im| and |K(ky) - K(-ky)| <= 2^-23 max|K|), as for the
// radially symmetric Lenia kernels; then fills `real` with rows ky = 0 .. N/2 of Re K in the
// kreal layout (types.h).
bool real_even_spectrum(const float* half, const float* nyq, uint32_t N, std::vector<float>& real) {
  const size_t H = N / 2;
  double mx = 0.0;
  for (size_t i = 0; i < N * H; ++i) mx = std::max(mx, std::hypot((double)half[2 * i], (double)half[2 * i + 1]));
  for (size_t y = 0; y < N; ++y) mx = std::max(mx, std::hypot((double)nyq[2 * y], (double)nyq[2 * y + 1]));
  const double tol = std::ldexp(mx, -23);
  for (size_t y = 0; y < N; ++y) {
    const size_t ym = (N - y) & (N - 1);
    for (size_t x = 0; x < H; ++x) {
      const float* a = half + 2 * (y * H + x);
      const float* b = half + 2 * (ym * H + x);
      if (std::fabs((double)a[1]) > tol || std::fabs((double)a[0] - (double)b[0]) > tol) return false;
    }
  }
  real.resize((H + 1) * H);
  for (size_t y = 0; y <= H; ++y)
    for (size_t x = 0; x < H; ++x) real[kreal_index((int)N, (int)y, (int)x)] = half[2 * (y * H + x)];
  return true;
}

}  // namespace cn

using namespace cn;

namespace {

struct Creature {
  uint32_t N = 0, C = 0, K = 0;
  CreatureDev dev{};       // host copy of the device descriptor
  int cls = 0, idx = 0;    // size class and position in its descriptor array
  std::vector<void*> allocs;
  bool adj = false;        // adjoint buffers allocated
  int islots = 1;          // intermediate slots of A (and of u once allocated)
  bool has_target = false;
  // Per kernel: support radius of its column profiles for the segmented column pass (-1: the
  // kernel has none within kSegRcap, or its spectrum is not real and even).
  std::vector<int> seg_r;
  size_t plane() const { return (size_t)N * N; }
  size_t slot_elems() const { return (size_t)C * N * N; }
};

struct SizeClass {
  int N = 0;
  std::vector<int> members;
  CreatureDev* d_descs = nullptr;
  int max_c = 0, max_k = 0;
  int fwd_e = 0, inv_e = 0, col_e = 0, col_tc = 0, diff_th = 0;  // launch geometry (0: default)
  bool col_split = false;
  bool col_full = false;  // full-length column transforms even for creatures with segments
};

// A contiguous range of one size class's creatures (descriptors d_descs[first .. first+count)):
// the unit that the step and rollout graphs put on a branch of their own.  Groups share no data,
// so with more than one group each runs on a branch stream, forked from and joined back into the
// engine's stream, and the groups' passes overlap (a column pass beside row or stencil passes).
struct LaunchGroup {
  int cls = 0, first = 0, count = 0;
};

// Captured rollout graphs of one (T, checkpoint stride) plan.
struct Rollout {
  uint32_t T = 0, ckpt = 0;
  hipGraphExec_t graph = nullptr;   // forward + loss + backward (targets)
  hipGraphExec_t fwd = nullptr;     // forward only
  int final_slot = 0;
  int grad_slot = 0;                // gradient slot holding dL/dstate0 after the graph's backward
};

// What the last rollout calls left on the device.  Kept apart from the graphs: dropping the
// graphs (a new launch plan) does not move any buffer, so these stay valid.
struct RolloutResult {
  uint32_t T = 0, ckpt = 0;         // plan of the last successful forward
  int final_slot = 0;               // state slot holding x_T of that forward
  bool fwd_valid = false;           // its checkpoints and x_T are intact
  int grad_slot = 0;                // gradient slot holding dL/dstate0 of the last backward
  bool grads_valid = false;         // a backward completed (gradients readable)
  bool loss_valid = false;          // a cn_rollout_grad completed (loss partials readable)
};

}  // namespace

struct cn_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<Creature> cr;
  std::vector<SizeClass> cls;
  int nslots = 2;
  int cur = 0;              // current-state slot (0 or 1)
  uint64_t steps = 0;
  bool have_prev = false;   // slot cur ^ 1 holds the input of the most recent step
  uint64_t bytes = 0;
  float2* tw = nullptr;
  double2* mass_partial = nullptr;
  int mass_nblk = kMassBlocks;
  unsigned long long* stats_partial = nullptr;  // frame statistics (kStatsBlocks x 6), on first use
  std::vector<void*> allocs;
  hipGraphExec_t step_graph[2] = {nullptr, nullptr};
  hipGraphExec_t render_graph[2] = {nullptr, nullptr};  // step + render epilogue
  // run_steps graphs of 8, 4 and 2 steps (kRunLen[r]), per ping-pong parity
  hipGraphExec_t multi_graph[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  RingParams* gen_rings = nullptr;  // device generator parameters, kMaxKernels entries
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  hipEvent_t join[2] = {nullptr, nullptr};  // cross-stream ordering (cn_iterate_on_stream)
  // Launch groups (see LaunchGroup) and, with more than one, a branch stream per group and the
  // fork/join events of the branches.
  std::vector<LaunchGroup> grp;
  // Backward pipelines (allocated with the first rollout): per group a recompute stream beside
  // the group's stream and the events that order the two (see record_backward).
  std::vector<hipStream_t> rstream;
  std::vector<std::array<hipEvent_t, 5>> pev;  // [fork, recompute done, adjoint done (2 parities), join]
  std::vector<hipStream_t> bstream;
  std::vector<hipEvent_t> bev;  // [group] join events, [groups] the fork event
  Rollout ro;               // last captured rollout plan
  RolloutResult res;        // results of the last rollout forward / backward
  // cn_adam_step writes (mu, sigma, inv_s2, w) into the device descriptors: until they are read
  // back, the host copies of those fields are stale (refresh_host_growth).
  bool growth_on_device = false;
  size_t nparams = 0;       // sum over creatures of 3K (batch-wide parameter arrays)
  float* grad_gather = nullptr;  // device staging of every creature's parameter gradients
};

namespace {

// Zeroed device allocation owned by `owner`.  On failure nothing is kept or counted and *p is
// left null.
cn_status dalloc(cn_engine* e, std::vector<void*>& owner, void** p, size_t bytes) {
  *p = nullptr;
  if (inject_failure()) return set_error(CN_ERR_HIP, "hipMalloc: injected failure");
  void* q = nullptr;
  CN_HIP(hipMalloc(&q, bytes));
  cn_status st;
  if (inject_failure()) {
    (void)hipFree(q);
    return set_error(CN_ERR_HIP, "hipMemsetAsync: injected failure");
  }
  if (hip_failed(hipMemsetAsync(q, 0, bytes, e->stream), "hipMemsetAsync", &st)) {
    (void)hipFree(q);
    return st;
  }
  owner.push_back(q);
  e->bytes += bytes;
  *p = q;
  return CN_OK;
}

// Releases an allocation made by dalloc (`bytes` is its size); a null pointer is ignored.
// hipFree of a live allocation has no recoverable failure, so its status is not propagated:
// the buffer leaves the owner list and the byte count together with the memory.
void dfree(cn_engine* e, std::vector<void*>& owner, void* p, size_t bytes) {
  if (!p) return;
  auto it = std::find(owner.begin(), owner.end(), p);
  if (it == owner.end()) return;
  owner.erase(it);
  e->bytes -= bytes;
  (void)hipFree(p);
}

void drop_graphs(cn_engine* e) {
  for (auto* gs : {e->step_graph, e->render_graph, e->multi_graph[0], e->multi_graph[1], e->multi_graph[2]})
    for (int j = 0; j < 2; ++j)
      if (gs[j]) {
        (void)hipGraphExecDestroy(gs[j]);
        gs[j] = nullptr;
      }
  if (e->ro.graph) (void)hipGraphExecDestroy(e->ro.graph);
  if (e->ro.fwd) (void)hipGraphExecDestroy(e->ro.fwd);
  e->ro = Rollout{};  // e->res is untouched: no buffer moves with the graphs
}

void release(cn_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graphs(e);
  for (auto& v : e->ev)
    if (v) (void)hipEventDestroy(v);
  for (auto& v : e->join)
    if (v) (void)hipEventDestroy(v);
  for (auto& q : e->bstream)
    if (q) (void)hipStreamSynchronize(q);
  for (auto& q : e->rstream)
    if (q) (void)hipStreamSynchronize(q);
  for (auto& a : e->pev)
    for (auto& v : a)
      if (v) (void)hipEventDestroy(v);
  for (auto& q : e->rstream)
    if (q) (void)hipStreamDestroy(q);
  for (auto& v : e->bev)
    if (v) (void)hipEventDestroy(v);
  for (auto& q : e->bstream)
    if (q) (void)hipStreamDestroy(q);
  for (auto& c : e->cr)
    for (void* p : c.allocs) (void)hipFree(p);
  for (void* p : e->allocs) (void)hipFree(p);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

// Reads the growth parameters an optimiser step left in the device descriptors back into the
// host copies (one copy per size class).  Synchronises.
cn_status refresh_host_growth(cn_engine* e) {
  if (!e->growth_on_device) return CN_OK;
  for (const SizeClass& k : e->cls) {
    std::vector<CreatureDev> tmp(k.members.size());
    CN_HIP(hipMemcpyAsync(tmp.data(), k.d_descs, tmp.size() * sizeof(CreatureDev), hipMemcpyDeviceToHost, e->stream));
    CN_HIP(hipStreamSynchronize(e->stream));
    for (size_t j = 0; j < k.members.size(); ++j) {
      CreatureDev& h = e->cr[k.members[j]].dev;
      const CreatureDev& g = tmp[j];
      std::memcpy(h.mu, g.mu, sizeof h.mu);
      std::memcpy(h.sigma, g.sigma, sizeof h.sigma);
      std::memcpy(h.inv_s2, g.inv_s2, sizeof h.inv_s2);
      std::memcpy(h.w, g.w, sizeof h.w);
    }
  }
  e->growth_on_device = false;
  return CN_OK;
}

cn_status upload_desc(cn_engine* e, int i) {
  // The host copy is about to overwrite the device descriptor: bring its growth fields up to
  // date first.
  const cn_status sr = refresh_host_growth(e);
  if (sr != CN_OK) return sr;
  Creature& c = e->cr[i];
  CN_HIP(hipMemcpyAsync(e->cls[c.cls].d_descs + c.idx, &c.dev, sizeof(CreatureDev), hipMemcpyHostToDevice,
                        e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  return CN_OK;
}

// Uploads creature i's host descriptor after a change to its buffers or flags (not to its
// growth parameters).  If that fails, the host copy goes back to `before` (what the device
// held; growth fields as the upload's read-back left them) and, unless the failure came before
// the copy was issued, is copied to the device again outside the checked wrapper, since the
// failed copy may still land.  *restored tells whether the device copy is then known to equal
// the host's: only then may the caller free buffers that the changed descriptor named.
cn_status commit_desc(cn_engine* e, int i, const CreatureDev& before, bool* restored) {
  const cn_status st = upload_desc(e, i);
  *restored = true;
  if (st == CN_OK) return CN_OK;
  Creature& c = e->cr[i];
  const CreatureDev now = c.dev;
  c.dev = before;
  std::memcpy(c.dev.mu, now.mu, sizeof now.mu);
  std::memcpy(c.dev.sigma, now.sigma, sizeof now.sigma);
  std::memcpy(c.dev.inv_s2, now.inv_s2, sizeof now.inv_s2);
  std::memcpy(c.dev.w, now.w, sizeof now.w);
  if (!e->growth_on_device)  // else the failure was the growth read-back: nothing was copied
    *restored = hipMemcpyAsync(e->cls[c.cls].d_descs + c.idx, &c.dev, sizeof(CreatureDev), hipMemcpyHostToDevice,
                               e->stream) == hipSuccess &&
                hipStreamSynchronize(e->stream) == hipSuccess;
  return st;
}

// Column-profile support radii and segment tables of kernels `ks` of creature i (after their
// spectra changed), then the creature's segment geometry: seg_g segments of M = seg_m(N) rows
// overlapping by 2 seg_r, seg_r the largest radius, when every kernel has one (else seg_g = 0,
// full-length column transforms).  Kernels keep radius -1 when anything fails.  Launch graphs
// bake the segment count and the choice of column pass in, so a change drops them.
cn_status build_segments(cn_engine* e, int i, const std::vector<int>& ks) {
  Creature& c = e->cr[i];
  const int N = (int)c.N, M = seg_m(N), W = N / 2 + 1, RC = kSegRcap;
  std::vector<int> todo;
  for (int k : ks) {
    c.seg_r[k] = -1;
    if (M > 0 && c.dev.ksym[k]) todo.push_back(k);
  }
  cn_status st = CN_OK;
  auto ok = [&](hipError_t he, const char* what) {
    if (st != CN_OK) return false;
    if (inject_failure()) {
      st = set_error(CN_ERR_HIP, "%s: injected failure", what);
      return false;
    }
    return !hip_failed(he, what, &st);
  };
  if (!todo.empty()) {
    const int nk = (int)todo.size();
    std::vector<void*> owner;
    std::vector<std::pair<void*, size_t>> tmp;
    auto talloc = [&](size_t bytes) -> void* {
      void* q = nullptr;
      if (st == CN_OK && (st = dalloc(e, owner, &q, bytes)) == CN_OK) tmp.emplace_back(q, bytes);
      return q;
    };
    const size_t sb = (size_t)nk * (RC + 2) * 8;
    double* hbuf = (double*)talloc((size_t)nk * W * (RC + 1) * 8);
    double* etail = (double*)talloc((size_t)nk * W * (RC + 2) * 8);
    double* esum = (double*)talloc(sb);
    int* dks = (int*)talloc((size_t)nk * 4);
    int* drs = (int*)talloc((size_t)nk * 4);
    std::vector<double> es((size_t)nk * (RC + 2));
    const CreatureDev* d = e->cls[c.cls].d_descs + c.idx;
    if (st == CN_OK && ok(hipMemcpyAsync(dks, todo.data(), (size_t)nk * 4, hipMemcpyHostToDevice, e->stream),
                          "hipMemcpyAsync") &&
        ok(launch_seg_profile(N, d, dks, nk, hbuf, etail, e->stream), "segment profiles") &&
        ok(launch_seg_tail_sum(N, etail, nk, esum, e->stream), "segment tail sums") &&
        ok(hipMemcpyAsync(es.data(), esum, sb, hipMemcpyDeviceToHost, e->stream), "hipMemcpyAsync") &&
        ok(hipStreamSynchronize(e->stream), "hipStreamSynchronize")) {
      // Support: the smallest radius whose tail energy is within (2^-22)^2 of the total.
      const double tau2 = std::ldexp(1.0, -44);
      std::vector<int> rad(nk, -1);
      for (int kk = 0; kk < nk; ++kk) {
        const double* ek = es.data() + (size_t)kk * (RC + 2);
        for (int r = 0; r <= RC && rad[kk] < 0; ++r)
          if (ek[r + 1] <= tau2 * ek[0]) rad[kk] = r;
      }
      bool any = false;
      for (int r : rad) any = any || r >= 0;
      if (any && ok(hipMemcpyAsync(drs, rad.data(), (size_t)nk * 4, hipMemcpyHostToDevice, e->stream),
                    "hipMemcpyAsync") &&
          ok(launch_seg_table(N, hbuf, dks, drs, nk, const_cast<float*>(c.dev.kseg), e->stream), "segment tables") &&
          ok(hipStreamSynchronize(e->stream), "hipStreamSynchronize")) {
        for (int kk = 0; kk < nk; ++kk) c.seg_r[todo[kk]] = rad[kk];
      }
    }
    (void)hipStreamSynchronize(e->stream);
    for (auto& q : tmp) dfree(e, owner, q.first, q.second);
  }
  int R = -1;
  bool all = M > 0;
  for (int r : c.seg_r) {
    all = all && r >= 0;
    R = std::max(R, r);
  }
  all = all && R <= seg_rmax(N);
  const int g = all ? (N + (M - 2 * R) - 1) / (M - 2 * R) : 0, r = all ? R : 0;
  if (g != c.dev.seg_g || r != c.dev.seg_r) {
    const CreatureDev before = c.dev;
    c.dev.seg_g = g;
    c.dev.seg_r = r;
    drop_graphs(e);
    bool restored = false;
    const cn_status cs = commit_desc(e, i, before, &restored);
    if (st == CN_OK) st = cs;
  }
  return st;
}

// Segmented column pass of creatures members[first .. first+count) of a class: the largest
// seg_g (0: none has segments, or the class's plan asks for full-length column transforms)
// and whether some of them have none (the full-length pass then runs for those).
void set_seg_grid(const cn_engine* e, const SizeClass& k, int first, int count, StepLaunch& sl) {
  sl.seg_g = 0;
  sl.seg_mixed = false;
  if (k.col_full) return;
  for (int j = first; j < first + count; ++j) {
    const int cg = e->cr[k.members[j]].dev.seg_g;
    if (cg > 0)
      sl.seg_g = std::max(sl.seg_g, cg);
    else
      sl.seg_mixed = true;
  }
  if (sl.seg_g == 0) sl.seg_mixed = false;
}

StepLaunch class_launch(cn_engine* e, const SizeClass& k, int sin, int sout) {
  StepLaunch sl;
  sl.descs = k.d_descs;
  sl.count = (int)k.members.size();
  sl.sin = sin;
  sl.sout = sout;
  sl.max_c = k.max_c;
  sl.max_k = k.max_k;
  sl.tw = e->tw;
  sl.fwd_e = k.fwd_e;
  sl.inv_e = k.inv_e;
  sl.col_e = k.col_e;
  sl.col_tc = k.col_tc;
  sl.diff_th = k.diff_th;
  sl.col_split = k.col_split;
  set_seg_grid(e, k, 0, (int)k.members.size(), sl);
  return sl;
}

// The launch of group q's creatures: its class's geometry over its range of descriptors.
StepLaunch group_launch(cn_engine* e, int q, int sin, int sout) {
  const LaunchGroup& g = e->grp[q];
  StepLaunch sl = class_launch(e, e->cls[g.cls], sin, sout);
  sl.descs = e->cls[g.cls].d_descs + g.first;
  sl.count = g.count;
  set_seg_grid(e, e->cls[g.cls], g.first, g.count, sl);
  return sl;
}
hipStream_t group_stream(const cn_engine* e, int q) { return e->bstream.empty() ? e->stream : e->bstream[q]; }
int group_n(const cn_engine* e, int q) { return e->cls[e->grp[q].cls].N; }

// Branches: every group's stream waits for the work queued so far on the engine's stream
// (fork), and the engine's stream waits for every group's work (join).  No-ops with one group.
hipError_t fork_groups(cn_engine* e) {
  if (e->bstream.empty()) return hipSuccess;
  const size_t ng = e->grp.size();
  hipError_t err = hipEventRecord(e->bev[ng], e->stream);
  for (size_t q = 0; q < ng && err == hipSuccess; ++q) err = hipStreamWaitEvent(e->bstream[q], e->bev[ng], 0);
  return err;
}
hipError_t join_groups(cn_engine* e) {
  if (e->bstream.empty()) return hipSuccess;
  hipError_t err = hipSuccess;
  for (size_t q = 0; q < e->grp.size() && err == hipSuccess; ++q) {
    err = hipEventRecord(e->bev[q], e->bstream[q]);
    if (err == hipSuccess) err = hipStreamWaitEvent(e->stream, e->bev[q], 0);
  }
  return err;
}

// `steps` steps of group q on its stream, the first from state slot `sin` into `sout`, then
// alternating between the two (ping-pong).
hipError_t launch_group_runs(cn_engine* e, int q, int sin, int sout, int steps, bool render = false) {
  for (int j = 0; j < steps; ++j) {
    StepLaunch sl = (j & 1) ? group_launch(e, q, sout, sin) : group_launch(e, q, sin, sout);
    sl.render = render;
    const hipError_t err = launch_step(group_n(e, q), sl, group_stream(e, q));
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

// `steps` steps of every creature, each group on its branch.
hipError_t launch_runs(cn_engine* e, int sin, int sout, int steps, bool render = false) {
  hipError_t err = fork_groups(e);
  for (size_t q = 0; q < e->grp.size() && err == hipSuccess; ++q)
    err = launch_group_runs(e, (int)q, sin, sout, steps, render);
  return err == hipSuccess ? join_groups(e) : err;
}

hipError_t launch_all_steps(cn_engine* e, int sin, int sout, bool render = false) {
  return launch_runs(e, sin, sout, 1, render);
}

// A step of group q that keeps its intermediates (u with write_u, A) in slot `islot`;
// skip_diffusion stops before the new state.
hipError_t launch_group_keep(cn_engine* e, int q, int sin, int sout, bool skip_diffusion, int islot) {
  StepLaunch sl = group_launch(e, q, sin, sout);
  sl.write_u = true;
  sl.skip_diffusion = skip_diffusion;
  sl.islot = islot;
  return launch_step(group_n(e, q), sl, group_stream(e, q));
}

// Adjoint step of group q from state slot `sin` and the intermediates in slot `islot`
// (recomputed first when `recompute`).
hipError_t launch_group_adjoint(cn_engine* e, int q, int sin, int gin, int gout, int islot, bool recompute) {
  StepLaunch sl = group_launch(e, q, sin, sin);
  sl.islot = islot;
  sl.adj_recompute = recompute;
  return launch_adjoint(group_n(e, q), sl, gin, gout, group_stream(e, q));
}

// Captures `fn` (a sequence of launches on e->stream) into an executable graph.
template <class F>
cn_status capture(cn_engine* e, hipGraphExec_t* out, F&& fn) {
  hipGraph_t g = nullptr;
  CN_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  const hipError_t le = fn();
  const hipError_t ce = hipStreamEndCapture(e->stream, &g);
  cn_status st;
  if (hip_failed(le, "kernel launch (capture)", &st)) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (hip_failed(ce, "hipStreamEndCapture", &st)) return st;
  const hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (hip_failed(ie, "hipGraphInstantiate", &st)) return st;
  return CN_OK;
}

constexpr int kRunLen[3] = {8, 4, 2};  // steps per run_steps graph (even: the parity returns)

// One step as a graph launch on `on` (the engine's stream unless a caller's stream is given:
// then the caller's stream first waits for the engine's queued work and the engine's stream
// afterwards waits for the step, so both stay ordered without a host wait).
cn_status enqueue_step(cn_engine* e, bool render = false, hipStream_t on = nullptr, bool foreign = false) {
  const int cur = e->cur;
  hipGraphExec_t* g = render ? &e->render_graph[cur] : &e->step_graph[cur];
  if (!*g) {
    const cn_status st = capture(e, g, [&] { return launch_all_steps(e, cur, cur ^ 1, render); });
    if (st != CN_OK) return st;
  }
  if (foreign) {
    CN_HIP(hipEventRecord(e->join[0], e->stream));
    CN_HIP(hipStreamWaitEvent(on, e->join[0], 0));
    CN_HIP(hipGraphLaunch(*g, on));
    CN_HIP(hipEventRecord(e->join[1], on));
    CN_HIP(hipStreamWaitEvent(e->stream, e->join[1], 0));
  } else {
    CN_HIP(hipGraphLaunch(*g, e->stream));
  }
  e->cur ^= 1;
  e->steps++;
  e->have_prev = true;
  e->res.fwd_valid = false;
  return CN_OK;
}

cn_status check_creature(const cn_engine* e, uint32_t i) {
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  if (i >= e->cr.size()) return set_error(CN_ERR_INVALID_ARG, "creature %u out of range (%zu)", i, e->cr.size());
  return CN_OK;
}

// Grows every creature's state-slot array to `want` slots (slots 0/1 are preserved).
cn_status ensure_slots(cn_engine* e, int want) {
  if (want <= e->nslots) return CN_OK;
  CN_HIP(hipStreamSynchronize(e->stream));
  for (size_t i = 0; i < e->cr.size(); ++i) {
    Creature& c = e->cr[i];
    // A creature already grown by an earlier call that failed part-way keeps its buffer.
    if (c.dev.nslots >= want) continue;
    void* p = nullptr;
    const size_t bytes = (size_t)want * c.slot_elems() * 4;
    cn_status st = dalloc(e, c.allocs, &p, bytes);
    if (st != CN_OK) return st;
    st = [&]() -> cn_status {
      CN_HIP(hipMemcpyAsync(p, c.dev.state, 2 * c.slot_elems() * 4, hipMemcpyDeviceToDevice, e->stream));
      CN_HIP(hipStreamSynchronize(e->stream));
      return CN_OK;
    }();
    if (st != CN_OK) {  // no descriptor names the new buffer yet
      dfree(e, c.allocs, p, bytes);
      return st;
    }
    // New buffer into the descriptor first; the old one is freed only once the device copy
    // of the descriptor no longer refers to it.
    const CreatureDev before = c.dev;
    c.dev.state = (float*)p;
    c.dev.nslots = want;
    bool restored = false;
    if ((st = commit_desc(e, (int)i, before, &restored)) != CN_OK) {
      if (restored) dfree(e, c.allocs, p, bytes);
      return st;
    }
    dfree(e, c.allocs, before.state, (size_t)before.nslots * c.slot_elems() * 4);
  }
  e->nslots = want;
  return CN_OK;
}

cn_status ensure_adjoint(cn_engine* e) {
  for (size_t i = 0; i < e->cr.size(); ++i) {
    Creature& c = e->cr[i];
    if (c.adj) continue;
    const size_t P = c.plane(), C = c.C, K = c.K, H = c.N / 2;
    CreatureDev& d = c.dev;
    void* p = nullptr;
    cn_status st;
    if (!d.u) {
      if ((st = dalloc(e, c.allocs, &p, K * P * 4)) != CN_OK) return st;
      d.u = (float*)p;
    }
    // Each buffer is allocated once: a retry after a failure part-way keeps the ones it has.
    auto need = [&](auto*& field, size_t bytes) -> cn_status {
      if (field) return CN_OK;
      void* q = nullptr;
      const cn_status s2 = dalloc(e, c.allocs, &q, bytes);
      if (s2 == CN_OK) field = (std::remove_reference_t<decltype(field)>)q;
      return s2;
    };
    if ((st = need(d.grad, 2 * C * P * 4)) != CN_OK) return st;
    if ((st = need(d.gaff, C * P * 4)) != CN_OK) return st;
    if ((st = need(d.dsd, C * P * 4)) != CN_OK) return st;
    if ((st = need(d.specg, K * c.N * H * 8)) != CN_OK) return st;
    if ((st = need(d.adjspec, C * c.N * H * 8)) != CN_OK) return st;
    if ((st = need(d.partial, (size_t)row_blocks((int)c.N) * K * 3 * 4)) != CN_OK) return st;
    if ((st = need(d.gparam, 6 * K * 4)) != CN_OK) return st;
    if ((st = need(d.target, C * P * 4)) != CN_OK) return st;
    if ((st = need(d.lossp, (size_t)kLossPartials * 4)) != CN_OK) return st;
    // The buffers stay with the creature whatever happens to the upload; a retry uploads again.
    if ((st = upload_desc(e, (int)i)) != CN_OK) return st;
    c.adj = true;
  }
  return CN_OK;
}

// Grows every creature's intermediate arrays (A: C planes, u: K planes per slot) to `want`
// slots; their contents are recomputed before every use, so nothing is copied.
cn_status ensure_islots(cn_engine* e, int want) {
  bool sync = false;
  for (size_t i = 0; i < e->cr.size(); ++i) {
    Creature& c = e->cr[i];
    if (c.islots >= want && c.dev.u) continue;
    if (!sync) {
      CN_HIP(hipStreamSynchronize(e->stream));
      sync = true;
    }
    const int n = std::max(want, c.islots);
    const size_t a_bytes = (size_t)n * c.C * c.plane() * 4, u_bytes = (size_t)n * c.K * c.plane() * 4;
    // Both new arrays first, then the descriptor, then the old arrays: a failure leaves the
    // creature on its old, consistent buffers.
    void *pa = nullptr, *pu = nullptr;
    cn_status st = dalloc(e, c.allocs, &pa, a_bytes);
    if (st != CN_OK) return st;
    // On a failure the new arrays are released again, so a retry neither keeps nor counts
    // them twice.
    if ((st = dalloc(e, c.allocs, &pu, u_bytes)) != CN_OK) {
      dfree(e, c.allocs, pa, a_bytes);
      return st;
    }
    const CreatureDev before = c.dev;
    const size_t old_bytes[2] = {(size_t)c.islots * c.C * c.plane() * 4,
                                 c.dev.u ? (size_t)c.islots * c.K * c.plane() * 4 : 0};
    c.dev.A = (float*)pa;
    c.dev.u = (float*)pu;
    bool restored = false;
    if ((st = commit_desc(e, (int)i, before, &restored)) != CN_OK) {
      if (restored) {
        dfree(e, c.allocs, pa, a_bytes);
        dfree(e, c.allocs, pu, u_bytes);
      }
      return st;
    }
    c.islots = n;
    dfree(e, c.allocs, before.A, old_bytes[0]);
    dfree(e, c.allocs, before.u, old_bytes[1]);
  }
  return CN_OK;
}

// Slot plan of a rollout of T steps with a checkpoint every `ck` steps, starting from slot
// `start` (0 or 1, left unmodified):
//   ck[0] = start; ck[j] (1 <= j < nck) = slot 2 + j - 1; forward scratch X0, X1;
//   recompute buffers seg[i] (1 <= i < ck); final state.
// State slots of a rollout: 0/1 ping-pong (the start is slot 0), checkpoints 2.. for segments
// 1.., two scratch slots x0, x0 + 1 for the forward, and two sets of ck - 1 segment slots (the
// states inside a segment) so that one segment can be recomputed while the adjoint still reads
// the states of the next one.  Segment j uses set par(j) = (nck - 1 - j) mod 2: the forward
// keeps the last segment's states in set 0.  fin holds the final state.
struct SlotPlan {
  int nck, x0, seg0, fin, seg1, total;
  int ckslot(int j, int start) const { return j == 0 ? start : 2 + j - 1; }
  int segslot(int i, int par = 0) const { return (par ? seg1 : seg0) + i - 1; }
  int par(int j) const { return (nck - 1 - j) & 1; }
};
SlotPlan plan_slots(uint32_t T, uint32_t ck) {
  if (ck == 0) ck = 1;  // callers validate T, ck > 0; never divide by zero
  SlotPlan p;
  p.nck = (int)((T + ck - 1) / ck);
  p.x0 = 2 + p.nck - 1;
  p.seg0 = p.x0 + 2;
  p.fin = p.seg0 + (int)ck - 1;
  p.seg1 = p.fin + 1;
  p.total = p.seg1 + (int)ck - 1;
  return p;
}

// The steps of the last segment keep their states (segment slots) and intermediates (slot
// t - tl), so a backward recorded right after it can skip that segment's recompute.
// Group q's forward (on its stream).
hipError_t record_forward(cn_engine* e, int q, const SlotPlan& p, uint32_t T, uint32_t ck, int start) {
  const uint32_t tl = (uint32_t)(p.nck - 1) * ck;  // first step of the last segment
  int in = start;
  for (uint32_t t = 0; t < T; ++t) {
    int out;
    if (t + 1 == T) out = p.fin;
    else if ((t + 1) % ck == 0) out = p.ckslot((int)((t + 1) / ck), start);
    else if (t + 1 > tl) out = p.segslot((int)(t + 1 - tl));
    else out = p.x0 + (int)(t % 2);
    const hipError_t err = t >= tl ? launch_group_keep(e, q, in, out, false, (int)(t - tl))
                                   : launch_group_runs(e, q, in, out, 1);
    if (err != hipSuccess) return err;
    in = out;
  }
  return hipSuccess;
}

// Backward of group q from gradient slot 0 (dL/dfinal); returns the slot holding dL/dstate0.
// last_kept: the last segment's states and intermediates are still those record_forward left.
//
// Segment j's recompute (from its checkpoint, keeping each step's states and intermediates) only
// needs the checkpoint, so it runs on the group's recompute stream while the adjoint, on the
// group's stream, still walks back through segment j + 1: segments alternate between two sets
// of segment-state and intermediate slots (par(j)), the adjoint of segment j waits for its
// recompute (event 1), and the recompute of segment j waits for the adjoint of segment j + 2,
// the last reader of the same set (events 2 + par, recorded after each segment's adjoint).  The
// adjoint steps themselves stay in order on one stream, so gradients are bitwise those of the
// sequential schedule.
//
// Only engines with a single launch group pipeline their backward (the recompute stream forks
// from the engine's own stream).  With several groups the recompute stream would fork from a
// group's branch stream, itself forked from the engine's stream; the one run of that graph (the
// 32-creature search) ended the process with a segmentation fault in the host runtime whose
// cause is not known, so those engines recompute on the group's stream, in sequence.
hipError_t record_backward(cn_engine* e, int q, const SlotPlan& p, uint32_t T, uint32_t ck, int start,
                           bool last_kept, int* gslot) {
  hipError_t err;
  const bool pipe = e->grp.size() == 1;
  const hipStream_t as = group_stream(e, q), rs = pipe ? e->rstream[q] : as;
  const auto& ev = e->pev[q];
  const StepLaunch gl = group_launch(e, q, 0, 1);
  if ((err = launch_zero_gparam(gl.descs, gl.count, as)) != hipSuccess) return err;
  if (pipe) {  // fork the recompute stream
    if ((err = hipEventRecord(ev[0], as)) != hipSuccess) return err;
    if ((err = hipStreamWaitEvent(rs, ev[0], 0)) != hipSuccess) return err;
  }
  hipError_t jerr = hipSuccess;
  auto join = [&] {  // the recompute stream back into the group's stream (also after an error)
    if (!pipe) return hipSuccess;
    hipError_t e1 = hipEventRecord(ev[4], rs);
    if (e1 == hipSuccess) e1 = hipStreamWaitEvent(as, ev[4], 0);
    return e1;
  };
  int g = 0;
  for (int j = p.nck - 1; j >= 0; --j) {
    const uint32_t t0 = (uint32_t)j * ck, t1 = std::min(T, t0 + ck);
    const int par = p.par(j), ib = par * (int)ck;
    auto state_slot = [&](uint32_t t) { return t == t0 ? p.ckslot(j, start) : p.segslot((int)(t - t0), par); };
    // Recompute the segment from its checkpoint, keeping each step's intermediates (u, A) in
    // intermediate slot ib + t - t0; the last step stops before its diffusion (its output is the
    // next checkpoint or the final state, both stored).
    if (!(last_kept && j == p.nck - 1)) {
      if (pipe && j + 2 < p.nck && (err = hipStreamWaitEvent(rs, ev[2 + par], 0)) != hipSuccess) break;
      for (uint32_t t = t0; t < t1 && err == hipSuccess; ++t) {
        const bool last = t + 1 == t1;
        StepLaunch sl = group_launch(e, q, state_slot(t), last ? state_slot(t) : state_slot(t + 1));
        sl.write_u = true;
        sl.skip_diffusion = last;
        sl.islot = ib + (int)(t - t0);
        err = launch_step(group_n(e, q), sl, rs);
      }
      if (err != hipSuccess) break;
      if (pipe && ((err = hipEventRecord(ev[1], rs)) != hipSuccess || (err = hipStreamWaitEvent(as, ev[1], 0)) != hipSuccess))
        break;
    }
    for (uint32_t t = t1; t-- > t0 && err == hipSuccess;) {
      err = launch_group_adjoint(e, q, state_slot(t), g, g ^ 1, ib + (int)(t - t0), false);
      g ^= 1;
    }
    if (err != hipSuccess) break;
    if (pipe && (err = hipEventRecord(ev[2 + par], as)) != hipSuccess) break;
  }
  jerr = join();
  if (err != hipSuccess) return err;
  if (jerr != hipSuccess) return jerr;
  *gslot = g;
  return hipSuccess;
}

// The recompute streams and pipeline events of every group (first rollout; kept on a failure
// part-way, so that a retry creates only what is missing).
cn_status ensure_pipeline(cn_engine* e) {
  const size_t ng = e->grp.size();
  if (e->rstream.size() != ng) e->rstream.resize(ng, nullptr);
  if (e->pev.size() != ng) e->pev.resize(ng, {nullptr, nullptr, nullptr, nullptr, nullptr});
  for (size_t q = 0; q < ng; ++q) {
    if (!e->rstream[q]) CN_HIP(hipStreamCreateWithFlags(&e->rstream[q], hipStreamNonBlocking));
    for (auto& v : e->pev[q])
      if (!v) CN_HIP(hipEventCreateWithFlags(&v, hipEventDisableTiming));
  }
  return CN_OK;
}

cn_status prepare_rollout(cn_engine* e, uint32_t T, uint32_t ck) {
  if (T == 0 || ck == 0) return set_error(CN_ERR_INVALID_ARG, "steps and checkpoint stride must be > 0");
  if (ck > T) ck = T;
  // The forward about to run overwrites the previous one's checkpoints (and growing the slot
  // arrays below keeps only slots 0 and 1): its results are gone whatever happens next.
  e->res.fwd_valid = false;
  cn_status st;
  if ((st = ensure_adjoint(e)) != CN_OK) return st;
  const SlotPlan p = plan_slots(T, ck);
  if ((st = ensure_slots(e, p.total)) != CN_OK) return st;
  if ((st = ensure_islots(e, 2 * (int)ck)) != CN_OK) return st;  // two sets (record_backward)
  if ((st = ensure_pipeline(e)) != CN_OK) return st;
  if (e->ro.T == T && e->ro.ckpt == ck && e->ro.graph && e->ro.fwd) return CN_OK;
  if (e->ro.graph) (void)hipGraphExecDestroy(e->ro.graph);
  if (e->ro.fwd) (void)hipGraphExecDestroy(e->ro.fwd);
  e->ro = Rollout{};
  // Graphs always start from slot 0: the engine moves its current state there first.
  // Each launch group records its whole forward (and loss and backward) on its own branch.
  st = capture(e, &e->ro.fwd, [&] {
    hipError_t err = fork_groups(e);
    for (size_t q = 0; q < e->grp.size() && err == hipSuccess; ++q) err = record_forward(e, (int)q, p, T, ck, 0);
    return err == hipSuccess ? join_groups(e) : err;
  });
  if (st != CN_OK) return st;
  int gslot = 0;
  st = capture(e, &e->ro.graph, [&] {
    hipError_t err = fork_groups(e);
    for (size_t q = 0; q < e->grp.size() && err == hipSuccess; ++q) {
      if ((err = record_forward(e, (int)q, p, T, ck, 0)) != hipSuccess) break;
      const StepLaunch gl = group_launch(e, (int)q, 0, 1);
      if ((err = launch_mse_grad(gl.descs, gl.count, p.fin, 0, group_stream(e, (int)q))) != hipSuccess) break;
      err = record_backward(e, (int)q, p, T, ck, 0, true, &gslot);
    }
    return err == hipSuccess ? join_groups(e) : err;
  });
  if (st != CN_OK) return st;
  e->ro.T = T;
  e->ro.ckpt = ck;
  e->ro.final_slot = p.fin;
  e->ro.grad_slot = gslot;
  return CN_OK;
}

// Rollouts start from slot 0; move the current state there if it is in slot 1.
cn_status normalise_current(cn_engine* e) {
  if (e->cur == 0) return CN_OK;
  for (auto& c : e->cr)
    CN_HIP(hipMemcpyAsync(c.dev.state, c.dev.state + c.slot_elems(), c.slot_elems() * 4, hipMemcpyDeviceToDevice,
                          e->stream));
  e->cur = 0;
  e->have_prev = false;
  return CN_OK;
}

// Records the plan of the rollout forward just enqueued (its checkpoints and x_T are valid
// until the state, the parameters or the kernels change, or another rollout runs).
void set_forward_result(cn_engine* e) {
  e->res.T = e->ro.T;
  e->res.ckpt = e->ro.ckpt;
  e->res.final_slot = e->ro.final_slot;
  e->res.fwd_valid = true;
}

cn_status create_engine(const cn_engine_desc* descs, uint32_t count, int device, cn_engine** out) {
  if (!descs || !out) return set_error(CN_ERR_INVALID_ARG, "null descs or out");
  *out = nullptr;
  if (count == 0) return set_error(CN_ERR_INVALID_ARG, "creature count must be > 0");
  // The creatures of a size class are the z dimension of its launch grids.
  if (count > kMaxCreatures)
    return set_error(CN_ERR_UNSUPPORTED, "creature count %u exceeds %u", count, kMaxCreatures);
  for (uint32_t i = 0; i < count; ++i) {
    const cn_engine_desc& d = descs[i];
    const cn_status st =
        validate_creature(d.n, d.num_channels, d.num_kernels, d.src, d.dst, d.mu, d.sigma, d.weights, d.temp);
    if (st != CN_OK) {
      if (count > 1) {
        char msg[kErrorTextSize];
        std::memcpy(msg, last_error_text(), sizeof msg);
        set_error(st, "creature %u: %s", i, msg);
      }
      return st;
    }
  }
  if (device < 0) return set_error(CN_ERR_INVALID_ARG, "device %d", device);

  cn_engine* e = new cn_engine();
  e->device = device;
  auto fail = [&](cn_status s) {
    release(e);
    return s;
  };
  cn_status st;
  if (inject_failure()) return fail(set_error(CN_ERR_HIP, "hipSetDevice: injected failure"));
  if (hip_failed(hipSetDevice(device), "hipSetDevice", &st)) return fail(st);
  if (hip_failed(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "hipStreamCreate", &st)) return fail(st);
  // Timing events only (their readers synchronise the stream before touching device data):
  // without the system-scope fence a record does not write back and invalidate the caches,
  // which would lengthen the timed pass and slow the one after it.
  for (auto& v : e->ev)
    if (hip_failed(hipEventCreateWithFlags(&v, hipEventDisableSystemFence), "hipEventCreate", &st)) return fail(st);
  for (auto& v : e->join)
    if (hip_failed(hipEventCreateWithFlags(&v, hipEventDisableTiming), "hipEventCreate", &st)) return fail(st);

  // size classes in order of first appearance
  e->cr.resize(count);
  for (uint32_t i = 0; i < count; ++i) {
    const cn_engine_desc& d = descs[i];
    Creature& c = e->cr[i];
    c.N = d.n;
    c.C = d.num_channels;
    c.K = d.num_kernels;
    int ci = -1;
    for (size_t j = 0; j < e->cls.size(); ++j)
      if (e->cls[j].N == (int)d.n) ci = (int)j;
    if (ci < 0) {
      e->cls.push_back(SizeClass{});
      ci = (int)e->cls.size() - 1;
      e->cls[ci].N = (int)d.n;
    }
    SizeClass& k = e->cls[ci];

    c.cls = ci;
    c.idx = (int)k.members.size();
    k.members.push_back((int)i);
    k.max_c = std::max(k.max_c, (int)c.C);
    k.max_k = std::max(k.max_k, (int)c.K);
  }
  for (uint32_t i = 0; i < count; ++i) {
    const cn_engine_desc& dd = descs[i];
    Creature& c = e->cr[i];
    const size_t N = c.N, C = c.C, K = c.K, P = N * N, H = N / 2;
    CreatureDev& d = c.dev;
    d.N = (int)N;
    void* p = nullptr;
    if ((st = dalloc(e, c.allocs, &p, 2 * C * P * 4)) != CN_OK) return fail(st);
    d.state = (float*)p;
    d.nslots = 2;
    if ((st = dalloc(e, c.allocs, &p, C * N * H * 8)) != CN_OK) return fail(st);
    d.spec = p;
    if ((st = dalloc(e, c.allocs, &p, K * N * H * 8)) != CN_OK) return fail(st);
    d.khalf = p;
    if ((st = dalloc(e, c.allocs, &p, K * N * 8)) != CN_OK) return fail(st);
    d.knyq = p;
    if ((st = dalloc(e, c.allocs, &p, K * (H + 1) * H * 4)) != CN_OK) return fail(st);
    d.kreal = (const float*)p;
    if (seg_m((int)N) > 0) {
      if ((st = dalloc(e, c.allocs, &p, K * (size_t)(seg_m((int)N) / 2 + 1) * (H + 1) * 4)) != CN_OK) return fail(st);
      d.kseg = (const float*)p;
    }
    c.seg_r.assign(K, -1);
    if ((st = dalloc(e, c.allocs, &p, K * N * H * 8)) != CN_OK) return fail(st);
    d.uhat = p;
    if ((st = dalloc(e, c.allocs, &p, C * P * 4)) != CN_OK) return fail(st);
    d.A = (float*)p;
    fill_desc_params(d, c.C, c.K, dd.src, dd.dst, dd.mu, dd.sigma, dd.weights, dd.temp);
    d.param_ofs = (int)e->nparams;
    e->nparams += 3 * (size_t)c.K;
  }
  for (SizeClass& k : e->cls) {
    void* p = nullptr;
    if ((st = dalloc(e, e->allocs, &p, k.members.size() * sizeof(CreatureDev))) != CN_OK) return fail(st);
    k.d_descs = (CreatureDev*)p;
  }
  // Launch groups: every size class of several creatures is split in two halves, each a branch
  // of the graphs, so that one half's column pass overlaps the other's row and stencil passes
  // (and other classes' passes).  Measured: 32 x 512^2 C5K25 search 12,500 -> 13,600
  // grad-steps/s (four quarters: 12,900); the uneven 256-creature batch 4.77 ms per step with
  // the classes in sequence, 4.59 with one branch per class, 4.45 with two.
  for (size_t j = 0; j < e->cls.size(); ++j) {
    const int m = (int)e->cls[j].members.size();
    const int parts = m >= 2 ? 2 : 1;
    for (int q = 0; q < parts; ++q) e->grp.push_back(LaunchGroup{(int)j, m * q / parts, m * (q + 1) / parts - m * q / parts});
  }
  if (e->grp.size() > 1) {  // branch streams (default priority) and their fork/join events
    e->bstream.assign(e->grp.size(), nullptr);
    e->bev.assign(e->grp.size() + 1, nullptr);
    for (auto& q : e->bstream) {
      if (inject_failure()) return fail(set_error(CN_ERR_HIP, "hipStreamCreate: injected failure"));
      if (hip_failed(hipStreamCreateWithFlags(&q, hipStreamNonBlocking), "hipStreamCreate", &st)) return fail(st);
    }
    for (auto& v : e->bev) {
      if (inject_failure()) return fail(set_error(CN_ERR_HIP, "hipEventCreate: injected failure"));
      if (hip_failed(hipEventCreateWithFlags(&v, hipEventDisableTiming), "hipEventCreate", &st)) return fail(st);
    }
  }
  for (uint32_t i = 0; i < count; ++i)
    if ((st = upload_desc(e, (int)i)) != CN_OK) return fail(st);
  void* p = nullptr;
  if ((st = dalloc(e, e->allocs, &p, (size_t)tw_table_size() * 8)) != CN_OK) return fail(st);
  e->tw = (float2*)p;
  const std::vector<float> tw = make_twiddles();
  if ((st = dalloc(e, e->allocs, &p, (size_t)count * kMaxChannels * e->mass_nblk * sizeof(double2))) != CN_OK)
    return fail(st);
  e->mass_partial = (double2*)p;
  if (hip_failed(hipMemcpyAsync(e->tw, tw.data(), tw.size() * 4, hipMemcpyHostToDevice, e->stream),
                 "hipMemcpy twiddles", &st))
    return fail(st);
  if (hip_failed(hipStreamSynchronize(e->stream), "hipStreamSynchronize", &st)) return fail(st);
  *out = e;
  return CN_OK;
}

float* cur_state(cn_engine* e, uint32_t i) {
  Creature& c = e->cr[i];
  return c.dev.state + (size_t)e->cur * c.slot_elems();
}

}  // namespace

extern "C" {

void cn_debug_fail_next_hip(int count) { g_fail_next = count > 0 ? count : 0; }

void cn_debug_fail_hip_at(int skip) { g_fail_at = skip >= 0 ? skip : -1; }

cn_status cn_engine_create(const cn_engine_desc* desc, cn_engine** out) {
  CN_TRY
  clear_error();
  if (!desc) return set_error(CN_ERR_INVALID_ARG, "null desc or out");
  return create_engine(desc, 1, desc->device, out);
  CN_CATCH
}

cn_status cn_batch_create(const cn_engine_desc* descs, uint32_t count, int device, cn_engine** out) {
  CN_TRY
  clear_error();
  return create_engine(descs, count, device, out);
  CN_CATCH
}

cn_status cn_engine_destroy(cn_engine* e) {
  CN_TRY
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  release(e);
  return CN_OK;
  CN_CATCH
}

cn_status cn_creature_set_kernel(cn_engine* e, uint32_t i, uint32_t k, const float* interleaved, size_t n_complex) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!interleaved) return set_error(CN_ERR_INVALID_ARG, "null argument");
  Creature& c = e->cr[i];
  if (k >= c.K) return set_error(CN_ERR_INVALID_ARG, "kernel index %u out of range (K = %u)", k, c.K);
  const size_t N = c.N, H = N / 2;
  if (n_complex != N * N)
    return set_error(CN_ERR_INVALID_ARG, "kernel spectrum has %zu values, expected %zu", n_complex, N * N);
  if (!finite_all(interleaved, 2 * n_complex)) return set_error(CN_ERR_INVALID_ARG, "kernel spectrum not finite");
  std::vector<float> half(N * H * 2), nyq(N * 2);
  project_kernel(interleaved, c.N, half.data(), nyq.data());
  e->res.fwd_valid = false;  // the next backward would recompute with the new spectrum
  std::vector<float> real;
  const bool sym = real_even_spectrum(half.data(), nyq.data(), c.N, real);
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipMemcpyAsync((char*)c.dev.khalf + (size_t)k * N * H * 8, half.data(), half.size() * 4,
                        hipMemcpyHostToDevice, e->stream));
  CN_HIP(hipMemcpyAsync((char*)c.dev.knyq + (size_t)k * N * 8, nyq.data(), nyq.size() * 4, hipMemcpyHostToDevice,
                        e->stream));
  if (sym)
    CN_HIP(hipMemcpyAsync((char*)c.dev.kreal + (size_t)k * (H + 1) * H * 4, real.data(), real.size() * 4,
                          hipMemcpyHostToDevice, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  if (c.dev.ksym[k] != (sym ? 1 : 0)) {
    const CreatureDev before = c.dev;
    c.dev.ksym[k] = sym ? 1 : 0;
    bool restored = false;
    if ((st = commit_desc(e, (int)i, before, &restored)) != CN_OK) return st;
  }
  return build_segments(e, (int)i, {(int)k});
  CN_CATCH
}

cn_status cn_set_kernel(cn_engine* e, uint32_t k, const float* interleaved, size_t n_complex) {
  CN_TRY
  return cn_creature_set_kernel(e, 0, k, interleaved, n_complex);
  CN_CATCH
}

cn_status cn_creature_upload_channel_f32(cn_engine* e, uint32_t i, uint32_t channel, const float* data, size_t n) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!data) return set_error(CN_ERR_INVALID_ARG, "null argument");
  Creature& c = e->cr[i];
  if (channel >= c.C) return set_error(CN_ERR_INVALID_ARG, "channel %u out of range (C = %u)", channel, c.C);
  const size_t P = c.plane();
  if (n != P) return set_error(CN_ERR_INVALID_ARG, "channel has %zu values, expected %zu", n, P);
  if (!finite_all(data, P)) return set_error(CN_ERR_INVALID_ARG, "channel values not finite");
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipMemcpyAsync(cur_state(e, i) + channel * P, data, P * 4, hipMemcpyHostToDevice, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  e->have_prev = false;
  e->res.fwd_valid = false;
  return CN_OK;
  CN_CATCH
}

cn_status cn_upload_channel_f32(cn_engine* e, uint32_t channel, const float* data, size_t n) {
  CN_TRY
  return cn_creature_upload_channel_f32(e, 0, channel, data, n);
  CN_CATCH
}

cn_status cn_upload_channel_f64(cn_engine* e, uint32_t channel, const double* data, size_t n) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, 0);
  if (st != CN_OK) return st;
  if (!data) return set_error(CN_ERR_INVALID_ARG, "null argument");
  const Creature& c = e->cr[0];
  if (channel >= c.C) return set_error(CN_ERR_INVALID_ARG, "channel %u out of range (C = %u)", channel, c.C);
  if (n != c.plane()) return set_error(CN_ERR_INVALID_ARG, "channel has %zu values, expected %zu", n, c.plane());
  std::vector<float> f(n);
  for (size_t j = 0; j < n; ++j) {
    if (!std::isfinite(data[j])) return set_error(CN_ERR_INVALID_ARG, "channel value %zu not finite", j);
    f[j] = (float)data[j];  // mod.rs:220
  }
  return cn_creature_upload_channel_f32(e, 0, channel, f.data(), n);
  CN_CATCH
}

cn_status cn_creature_set_growth(cn_engine* e, uint32_t i, const float* mu, const float* sigma, const float* w,
                                 size_t K) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  Creature& c = e->cr[i];
  if (K != c.K) return set_error(CN_ERR_INVALID_ARG, "expected %u growth parameters, got %zu", c.K, K);
  if ((st = validate_growth(c.K, mu, sigma, w)) != CN_OK) return st;
  // A backward must recompute its segments with the parameters of its own forward.
  e->res.fwd_valid = false;
  CN_HIP(hipSetDevice(e->device));
  if ((st = refresh_host_growth(e)) != CN_OK) return st;  // before this creature's copy changes
  fill_growth(c.dev, c.K, mu, sigma, w);
  return upload_desc(e, (int)i);
  CN_CATCH
}

// ---------------------------------------------------------------------------------------
// Device creature generator and spectrum readback
// ---------------------------------------------------------------------------------------
cn_status cn_creature_generate(cn_engine* e, uint32_t i, const cn_creature_params* p) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if ((st = check_creature_params(p)) != CN_OK) return st;
  Creature& c = e->cr[i];
  if (p->n != c.N || p->num_channels != c.C || p->num_kernels != c.K)
    return set_error(CN_ERR_INVALID_ARG, "params describe %u^2 C%u K%u, creature %u is %u^2 C%u K%u", p->n,
                     p->num_channels, p->num_kernels, i, c.N, c.C, c.K);
  for (uint32_t k = 0; k < c.K; ++k)
    if ((int)p->src[k] != c.dev.src_of[k] || (int)p->dst[k] != c.dev.dst_of[k])
      return set_error(CN_ERR_INVALID_ARG, "kernel %u: src/dst differ from the creature's", k);
  std::vector<float> mu(c.K), sg(c.K), w(c.K);
  generate_growth(p, mu.data(), sg.data(), w.data());
  if ((st = validate_growth(c.K, mu.data(), sg.data(), w.data())) != CN_OK) return st;
  std::vector<RingParams> rings(c.K);
  for (uint32_t k = 0; k < c.K; ++k) rings[k] = ring_params(p->seed, p->creature, k, p->radius[k]);
  CN_HIP(hipSetDevice(e->device));
  if (!e->gen_rings) {
    void* q = nullptr;
    if ((st = dalloc(e, e->allocs, &q, kMaxKernels * sizeof(RingParams))) != CN_OK) return st;
    e->gen_rings = (RingParams*)q;
  }
  if ((st = refresh_host_growth(e)) != CN_OK) return st;  // before this creature's copy changes
  fill_growth(c.dev, c.K, mu.data(), sg.data(), w.data());
  // A periodic kernel that depends on dx^2 + dy^2 only is even in x and y, so its spectrum is
  // real and even: every generated kernel uses the compact real copy.
  for (uint32_t k = 0; k < c.K; ++k) c.dev.ksym[k] = 1;
  if ((st = upload_desc(e, (int)i)) != CN_OK) return st;
  const CreatureDev* d = e->cls[c.cls].d_descs + c.idx;
  CN_HIP(hipMemcpyAsync(e->gen_rings, rings.data(), c.K * sizeof(RingParams), hipMemcpyHostToDevice, e->stream));
  CN_HIP(launch_generate((int)c.N, d, e->gen_rings, (int)c.K, e->tw, e->stream));
  const SeedPatch sp = seed_patch_of(c.N, p->patch);
  CN_HIP(launch_seed(d, e->cur, c.slot_elems(), sp.side, sp.origin, p->seed, p->creature, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  e->have_prev = false;
  e->res.fwd_valid = false;
  std::vector<int> every(c.K);
  for (uint32_t k = 0; k < c.K; ++k) every[k] = (int)k;
  return build_segments(e, (int)i, every);
  CN_CATCH
}

cn_status cn_creature_download_kernels(cn_engine* e, uint32_t i, float* out, size_t n_floats) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  const Creature& c = e->cr[i];
  const size_t N = c.N, H = N / 2, P = N * N;
  if (n_floats != 2 * c.K * P)
    return set_error(CN_ERR_INVALID_ARG, "output has %zu floats, expected %zu", n_floats, 2 * c.K * P);
  CN_HIP(hipSetDevice(e->device));
  std::vector<float> half(2 * c.K * N * H), nyq(2 * c.K * N);
  CN_HIP(hipMemcpyAsync(half.data(), c.dev.khalf, half.size() * 4, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipMemcpyAsync(nyq.data(), c.dev.knyq, nyq.size() * 4, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  const float sc = (float)P;  // undo the 1/N^2 fold (exact: power of two)
  for (size_t k = 0; k < c.K; ++k) {
    const float* hk = half.data() + 2 * k * N * H;
    const float* nk = nyq.data() + 2 * k * N;
    float* o = out + 2 * k * P;
    for (size_t y = 0; y < N; ++y)
      for (size_t x = 0; x < N; ++x) {
        float re, im;
        if (x < H) {
          re = hk[2 * (y * H + x)];
          im = hk[2 * (y * H + x) + 1];
        } else if (x == H) {
          re = nk[2 * y];
          im = nk[2 * y + 1];
        } else {  // Hermitian: K(y, x) = conj K(-y, -x)
          const size_t ym = (N - y) & (N - 1), xm = N - x;
          re = hk[2 * (ym * H + xm)];
          im = -hk[2 * (ym * H + xm) + 1];
        }
        o[2 * (y * N + x)] = re * sc;
        o[2 * (y * N + x) + 1] = im * sc;
      }
  }
  return CN_OK;
  CN_CATCH
}

cn_status cn_iterate(cn_engine* e) {
  CN_TRY
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  CN_HIP(hipSetDevice(e->device));
  return enqueue_step(e);
  CN_CATCH
}

cn_status cn_run_steps(cn_engine* e, uint64_t n) {
  CN_TRY
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  CN_HIP(hipSetDevice(e->device));
  // Runs of 8, 4 and 2 steps replay one graph holding all of them (an even count, so the
  // ping-pong parity is back where it started): n steps take n / 8 launches of the 8-step graph,
  // then at most one of the 4- and 2-step graphs and one single step, instead of one launch per
  // step; each launch saved is a launch boundary (about 4.5 us at 2048^2).  Every parity's
  // graphs are captured by the first call, whatever its step count, so that later calls (a
  // timed run after a short warm-up, say) only launch.
  for (int r = 0; r < 3; ++r)
    for (int par = 0; par < 2; ++par) {
      if (e->multi_graph[r][par]) continue;
      const cn_status st = capture(e, &e->multi_graph[r][par], [&] { return launch_runs(e, par, par ^ 1, kRunLen[r]); });
      if (st != CN_OK) return st;
    }
  for (int r = 0; r < 3; ++r) {
    const uint64_t len = (uint64_t)kRunLen[r];
    while (n >= len) {
      CN_HIP(hipGraphLaunch(e->multi_graph[r][e->cur], e->stream));
      e->steps += len;
      e->have_prev = true;
      e->res.fwd_valid = false;
      n -= len;
      if (r > 0) break;  // n < 2 len here
    }
  }
  if (n == 1) {
    const cn_status st = enqueue_step(e);
    if (st != CN_OK) return st;
  }
  return CN_OK;
  CN_CATCH
}

cn_status cn_sync(cn_engine* e) {
  CN_TRY
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipStreamSynchronize(e->stream));
  return CN_OK;
  CN_CATCH
}

cn_status cn_iterate_on_stream(cn_engine* e, void* stream) {
  CN_TRY
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  CN_HIP(hipSetDevice(e->device));
  const hipStream_t s = (hipStream_t)stream;
  return enqueue_step(e, false, s, s != e->stream);
  CN_CATCH
}

cn_status cn_engine_stream(const cn_engine* e, void** stream) {
  CN_TRY
  clear_error();
  if (!e || !stream) return set_error(CN_ERR_INVALID_ARG, "null argument");
  *stream = (void*)e->stream;
  return CN_OK;
  CN_CATCH
}

cn_status cn_creature_state_ptr(cn_engine* e, uint32_t i, float** d_state, size_t* n) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!d_state) return set_error(CN_ERR_INVALID_ARG, "null argument");
  *d_state = cur_state(e, i);
  if (n) *n = e->cr[i].slot_elems();
  return CN_OK;
  CN_CATCH
}

// Allocates every creature's packed frame (zeroed, so byte 3 of each pixel stays 0).
static cn_status ensure_render(cn_engine* e) {
  for (size_t i = 0; i < e->cr.size(); ++i) {
    Creature& c = e->cr[i];
    if (c.dev.render) continue;
    void* p = nullptr;
    const cn_status st = dalloc(e, c.allocs, &p, c.plane() * 4);
    if (st != CN_OK) return st;
    const CreatureDev before = c.dev;
    c.dev.render = (unsigned int*)p;
    bool restored = false;
    const cn_status su = commit_desc(e, (int)i, before, &restored);
    if (su != CN_OK) {
      if (restored) dfree(e, c.allocs, p, c.plane() * 4);
      return su;
    }
  }
  return CN_OK;
}

cn_status cn_iterate_and_render(cn_engine* e) {
  CN_TRY
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  CN_HIP(hipSetDevice(e->device));
  const cn_status st = ensure_render(e);
  if (st != CN_OK) return st;
  return enqueue_step(e, true);
  CN_CATCH
}

cn_status cn_creature_download_render(cn_engine* e, uint32_t i, uint8_t* rgb, size_t n) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!rgb) return set_error(CN_ERR_INVALID_ARG, "null argument");
  const Creature& c = e->cr[i];
  const size_t P = c.plane();
  if (n != P * 3) return set_error(CN_ERR_INVALID_ARG, "output has %zu bytes, expected %zu", n, P * 3);
  if (!c.dev.render) {  // no frame rendered yet: the zero-initialised frame
    std::memset(rgb, 0, n);
    return CN_OK;
  }
  CN_HIP(hipSetDevice(e->device));
  std::vector<uint32_t> packed(P);
  CN_HIP(hipMemcpyAsync(packed.data(), c.dev.render, P * 4, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  for (size_t j = 0; j < P; ++j) {
    rgb[3 * j] = (uint8_t)(packed[j] >> 16);
    rgb[3 * j + 1] = (uint8_t)(packed[j] >> 8);
    rgb[3 * j + 2] = (uint8_t)packed[j];
  }
  return CN_OK;
  CN_CATCH
}

cn_status cn_download_render(cn_engine* e, uint8_t* rgb, size_t n) {
  CN_TRY
  return cn_creature_download_render(e, 0, rgb, n);
  CN_CATCH
}

cn_status cn_creature_render_features(cn_engine* e, uint32_t i, float* means, float* variances) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!means || !variances) return set_error(CN_ERR_INVALID_ARG, "null argument");
  const Creature& c = e->cr[i];
  const uint64_t cells = c.plane();
  uint64_t s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
  if (c.dev.render) {  // before the first rendered frame the frame is all zeros
    CN_HIP(hipSetDevice(e->device));
    if (!e->stats_partial) {
      void* p = nullptr;
      if ((st = dalloc(e, e->allocs, &p, (size_t)kStatsBlocks * 6 * 8)) != CN_OK) return st;
      e->stats_partial = (unsigned long long*)p;
    }
    CN_HIP(launch_render_stats(c.dev.render, cells, e->stats_partial, e->stream));
    std::vector<unsigned long long> part((size_t)kStatsBlocks * 6);
    CN_HIP(hipMemcpyAsync(part.data(), e->stats_partial, part.size() * 8, hipMemcpyDeviceToHost, e->stream));
    CN_HIP(hipStreamSynchronize(e->stream));
    for (int b = 0; b < kStatsBlocks; ++b)
      for (int ch = 0; ch < 3; ++ch) {
        s1[ch] += part[(size_t)6 * b + ch];
        s2[ch] += part[(size_t)6 * b + 3 + ch];
      }
  }
  features_from_sums(s1, s2, cells, means, variances);
  return CN_OK;
  CN_CATCH
}

cn_status cn_render_features(cn_engine* e, float* means, float* variances) {
  CN_TRY
  return cn_creature_render_features(e, 0, means, variances);
  CN_CATCH
}

// One directly launched step of every creature with events around every kernel of the first
// size class; kernel_ms receives [row forward, fused column, inverse row + growth, diffusion].
static cn_status timed_step(cn_engine* e, float kernel_ms[4]) {
  const int cur = e->cur;
  for (size_t j = 0; j < e->cls.size(); ++j) {
    StepLaunch sl = class_launch(e, e->cls[j], cur, cur ^ 1);
    if (j == 0)
      for (int q = 0; q < 5; ++q) sl.ev[q] = e->ev[q];
    CN_HIP(launch_step(e->cls[j].N, sl, e->stream));
  }
  CN_HIP(hipEventSynchronize(e->ev[4]));
  for (int q = 0; q < 4; ++q) CN_HIP(hipEventElapsedTime(&kernel_ms[q], e->ev[q], e->ev[q + 1]));
  CN_HIP(hipStreamSynchronize(e->stream));
  e->cur ^= 1;
  e->steps++;
  e->have_prev = true;
  e->res.fwd_valid = false;
  return CN_OK;
}

cn_status cn_timed_iterate(cn_engine* e, double ms[4]) {
  CN_TRY
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
  CN_CATCH
}

cn_status cn_profile_kernels(cn_engine* e, uint64_t steps, double* kernel_ms, size_t n) {
  CN_TRY
  clear_error();
  if (!e || !kernel_ms) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (n != 4) return set_error(CN_ERR_INVALID_ARG, "kernel_ms must have 4 entries");
  if (steps == 0) return set_error(CN_ERR_INVALID_ARG, "steps must be > 0");
  if (steps > 10000) return set_error(CN_ERR_INVALID_ARG, "steps must be <= 10000");
  CN_HIP(hipSetDevice(e->device));
  // 4 * steps direct steps queued back to back (no synchronisation between them, so the queue
  // never runs dry and every pass sees the caches a step leaves); step s carries two events
  // around pass s % 4 of the first size class only, so each pass is timed on `steps` steps with
  // no other event in its step.
  struct Events {
    std::vector<hipEvent_t> v;
    ~Events() {
      for (hipEvent_t x : v) (void)hipEventDestroy(x);
    }
  } evs;
  const uint64_t total = 4 * steps;
  for (uint64_t i = 0; i < 2 * total; ++i) {
    hipEvent_t x = nullptr;
    CN_HIP(hipEventCreateWithFlags(&x, hipEventDisableSystemFence));  // timing only, as e->ev
    evs.v.push_back(x);
  }
  for (uint64_t s = 0; s < total; ++s) {
    const int q = (int)(s % 4);
    for (size_t j = 0; j < e->cls.size(); ++j) {
      StepLaunch sl = class_launch(e, e->cls[j], e->cur, e->cur ^ 1);
      if (j == 0) {
        sl.ev[q] = evs.v[2 * s];
        sl.ev[q + 1] = evs.v[2 * s + 1];
      }
      CN_HIP(launch_step(e->cls[j].N, sl, e->stream));
    }
    e->cur ^= 1;
    e->steps++;
  }
  e->have_prev = true;
  e->res.fwd_valid = false;
  CN_HIP(hipStreamSynchronize(e->stream));
  double acc[4] = {0, 0, 0, 0};
  for (uint64_t s = 0; s < total; ++s) {
    float ms = 0.f;
    CN_HIP(hipEventElapsedTime(&ms, evs.v[2 * s], evs.v[2 * s + 1]));
    acc[s % 4] += ms;
  }
  for (int q = 0; q < 4; ++q) kernel_ms[q] = acc[q] / (double)steps;
  return CN_OK;
  CN_CATCH
}

cn_status cn_creature_download_channels(cn_engine* e, uint32_t i, float* out, size_t n) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  const size_t tot = e->cr[i].slot_elems();
  if (n != tot) return set_error(CN_ERR_INVALID_ARG, "output has %zu values, expected %zu", n, tot);
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipMemcpyAsync(out, cur_state(e, i), tot * 4, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  return CN_OK;
  CN_CATCH
}

cn_status cn_download_channels(cn_engine* e, float* out, size_t n) {
  CN_TRY
  return cn_creature_download_channels(e, 0, out, n);
  CN_CATCH
}

cn_status cn_download_conv_real(cn_engine* e, float* out, size_t n) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, 0);
  if (st != CN_OK) return st;
  if (!out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  Creature& c = e->cr[0];
  const size_t P = c.plane(), tot = (size_t)c.K * P;
  if (n != tot) return set_error(CN_ERR_INVALID_ARG, "output has %zu values, expected %zu", n, tot);
  if (!e->have_prev) {
    if (e->steps == 0) {
      std::memset(out, 0, tot * 4);
      return CN_OK;
    }
    return set_error(CN_ERR_STATE, "the input of the last step was overwritten since it ran");
  }
  CN_HIP(hipSetDevice(e->device));
  if (!c.dev.u) {
    void* p = nullptr;
    if ((st = dalloc(e, c.allocs, &p, tot * 4)) != CN_OK) return st;
    const CreatureDev before = c.dev;
    c.dev.u = (float*)p;
    bool restored = false;
    if ((st = commit_desc(e, 0, before, &restored)) != CN_OK) {
      if (restored) dfree(e, c.allocs, p, tot * 4);
      return st;
    }
  }
  // Recompute the convolution of the state the most recent step started from, for creature 0
  // alone: in a batch the other creatures may have no u buffer.
  const SizeClass& k = e->cls[c.cls];
  StepLaunch sl = class_launch(e, k, e->cur ^ 1, e->cur ^ 1);
  sl.descs = k.d_descs + c.idx;
  sl.count = 1;
  sl.write_u = true;
  sl.skip_diffusion = true;
  CN_HIP(launch_step(k.N, sl, e->stream));
  CN_HIP(hipMemcpyAsync(out, c.dev.u, tot * 4, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  const float scale = (float)P;  // unnormalised like the reference's conv buffer (exact: power of 2)
  for (size_t j = 0; j < tot; ++j) out[j] *= scale;
  return CN_OK;
  CN_CATCH
}

cn_status cn_creature_mass_fp64(cn_engine* e, uint32_t i, double* out, size_t n_channels) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  Creature& c = e->cr[i];
  if (n_channels != c.C) return set_error(CN_ERR_INVALID_ARG, "expected %u channels", c.C);
  CN_HIP(hipSetDevice(e->device));
  const SizeClass& k = e->cls[c.cls];
  CN_HIP(launch_mass(k.d_descs, (int)k.members.size(), e->cur, k.max_c, e->mass_partial, e->mass_nblk, e->stream));
  std::vector<double> part((size_t)kMaxChannels * e->mass_nblk * 2);
  CN_HIP(hipMemcpyAsync(part.data(), e->mass_partial + (size_t)c.idx * kMaxChannels * e->mass_nblk,
                        part.size() * 8, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  for (uint32_t ch = 0; ch < c.C; ++ch)
    out[ch] = mass_combine(reinterpret_cast<const double2*>(part.data()) + (size_t)ch * e->mass_nblk);
  return CN_OK;
  CN_CATCH
}

cn_status cn_mass_fp64(cn_engine* e, double* out, size_t n_channels) {
  CN_TRY
  return cn_creature_mass_fp64(e, 0, out, n_channels);
  CN_CATCH
}

cn_status cn_engine_get_info(const cn_engine* e, cn_engine_info* out) {
  CN_TRY
  clear_error();
  if (!e || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  out->n = e->cr[0].N;
  out->num_channels = e->cr[0].C;
  out->num_kernels = e->cr[0].K;
  out->steps_done = e->steps;
  out->device_bytes = e->bytes;
  out->compact_kernels = 0;
  for (uint32_t k = 0; k < e->cr[0].K; ++k) out->compact_kernels += e->cr[0].dev.ksym[k] ? 1u : 0u;
  out->segments = (uint32_t)e->cr[0].dev.seg_g;
  out->segment_radius = (uint32_t)e->cr[0].dev.seg_r;
  out->segment_length = e->cr[0].dev.seg_g > 0 ? (uint32_t)seg_m(e->cr[0].dev.N) : 0u;
  return CN_OK;
  CN_CATCH
}

// ---------------------------------------------------------------------------------------
// state <-> device pointers (caller-owned device memory; both calls synchronise)
// ---------------------------------------------------------------------------------------
cn_status cn_creature_state_from_device(cn_engine* e, uint32_t i, const float* d_src, size_t n) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!d_src) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (n != e->cr[i].slot_elems()) return set_error(CN_ERR_INVALID_ARG, "expected %zu values", e->cr[i].slot_elems());
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipMemcpyAsync(cur_state(e, i), d_src, n * 4, hipMemcpyDeviceToDevice, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  e->have_prev = false;
  e->res.fwd_valid = false;
  return CN_OK;
  CN_CATCH
}

cn_status cn_creature_state_to_device(cn_engine* e, uint32_t i, float* d_dst, size_t n) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!d_dst) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (n != e->cr[i].slot_elems()) return set_error(CN_ERR_INVALID_ARG, "expected %zu values", e->cr[i].slot_elems());
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipMemcpyAsync(d_dst, cur_state(e, i), n * 4, hipMemcpyDeviceToDevice, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  return CN_OK;
  CN_CATCH
}

// ---------------------------------------------------------------------------------------
// rollouts
// ---------------------------------------------------------------------------------------
cn_status cn_creature_set_target(cn_engine* e, uint32_t i, const float* target, size_t n) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!target) return set_error(CN_ERR_INVALID_ARG, "null argument");
  Creature& c = e->cr[i];
  if (n != c.slot_elems())
    return set_error(CN_ERR_INVALID_ARG, "target has %zu values, expected %zu", n, c.slot_elems());
  if (!finite_all(target, n)) return set_error(CN_ERR_INVALID_ARG, "target not finite");
  CN_HIP(hipSetDevice(e->device));
  if ((st = ensure_adjoint(e)) != CN_OK) return st;
  CN_HIP(hipMemcpyAsync(c.dev.target, target, n * 4, hipMemcpyHostToDevice, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  c.has_target = true;
  return CN_OK;
  CN_CATCH
}

cn_status cn_rollout_grad(cn_engine* e, uint32_t steps, uint32_t ckpt_every) {
  CN_TRY
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  for (size_t i = 0; i < e->cr.size(); ++i)
    if (!e->cr[i].has_target)
      return set_error(CN_ERR_STATE, "creature %zu has no target (cn_creature_set_target)", i);
  CN_HIP(hipSetDevice(e->device));
  cn_status st;
  if ((st = normalise_current(e)) != CN_OK) return st;
  if ((st = prepare_rollout(e, steps, ckpt_every)) != CN_OK) return st;
  e->res.grads_valid = false;
  e->res.loss_valid = false;
  CN_HIP(hipGraphLaunch(e->ro.graph, e->stream));
  set_forward_result(e);
  e->res.grad_slot = e->ro.grad_slot;
  e->res.grads_valid = true;
  e->res.loss_valid = true;
  return CN_OK;
  CN_CATCH
}

cn_status cn_rollout_forward(cn_engine* e, uint32_t steps, uint32_t ckpt_every) {
  CN_TRY
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  CN_HIP(hipSetDevice(e->device));
  cn_status st;
  if ((st = normalise_current(e)) != CN_OK) return st;
  if ((st = prepare_rollout(e, steps, ckpt_every)) != CN_OK) return st;
  CN_HIP(hipGraphLaunch(e->ro.fwd, e->stream));
  set_forward_result(e);
  return CN_OK;
  CN_CATCH
}

cn_status cn_creature_rollout_final(cn_engine* e, uint32_t i, float* out, size_t n, int device_pointer) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!e->res.fwd_valid) return set_error(CN_ERR_STATE, "no rollout forward since the last state change");
  Creature& c = e->cr[i];
  if (!out || n != c.slot_elems()) return set_error(CN_ERR_INVALID_ARG, "bad output buffer");
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipMemcpyAsync(out, c.dev.state + (size_t)e->res.final_slot * c.slot_elems(), n * 4,
                        device_pointer ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  return CN_OK;
  CN_CATCH
}

cn_status cn_rollout_backward(cn_engine* e, const float* const* d_grad_final, uint32_t count) {
  CN_TRY
  clear_error();
  if (!e || !d_grad_final) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (count != e->cr.size()) return set_error(CN_ERR_INVALID_ARG, "expected %zu gradient pointers", e->cr.size());
  if (!e->res.fwd_valid) return set_error(CN_ERR_STATE, "no rollout forward since the last state change");
  for (uint32_t i = 0; i < count; ++i)
    if (!d_grad_final[i]) return set_error(CN_ERR_INVALID_ARG, "null gradient pointer %u", i);
  e->res.grads_valid = false;
  CN_HIP(hipSetDevice(e->device));
  for (uint32_t i = 0; i < count; ++i) {
    const Creature& c = e->cr[i];
    CN_HIP(hipMemcpyAsync(c.dev.grad, d_grad_final[i], c.slot_elems() * 4, hipMemcpyDeviceToDevice, e->stream));
  }
  const SlotPlan p = plan_slots(e->res.T, e->res.ckpt);
  int gslot = 0;
  // Other calls may have used intermediate slot 0 since the forward: recompute every segment.
  CN_HIP(fork_groups(e));
  for (size_t q = 0; q < e->grp.size(); ++q) CN_HIP(record_backward(e, (int)q, p, e->res.T, e->res.ckpt, 0, false, &gslot));
  CN_HIP(join_groups(e));
  CN_HIP(hipStreamSynchronize(e->stream));
  e->res.grad_slot = gslot;
  e->res.grads_valid = true;
  return CN_OK;
  CN_CATCH
}

cn_status cn_creature_get_loss(cn_engine* e, uint32_t i, double* loss) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  if (!loss) return set_error(CN_ERR_INVALID_ARG, "null argument");
  Creature& c = e->cr[i];
  if (!c.adj || !e->res.loss_valid) return set_error(CN_ERR_STATE, "no cn_rollout_grad has completed");
  CN_HIP(hipSetDevice(e->device));
  std::vector<float> part(kLossPartials);
  CN_HIP(hipMemcpyAsync(part.data(), c.dev.lossp, part.size() * 4, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  double s = 0.0;
  for (float v : part) s += v;
  *loss = s / (double)c.slot_elems();
  return CN_OK;
  CN_CATCH
}

cn_status cn_creature_get_grads(cn_engine* e, uint32_t i, float* grad_state0, size_t n, float* grad_mu,
                                float* grad_sigma, float* grad_w, size_t K, int device_pointers) {
  CN_TRY
  clear_error();
  cn_status st = check_creature(e, i);
  if (st != CN_OK) return st;
  Creature& c = e->cr[i];
  if (!c.adj || !e->res.grads_valid) return set_error(CN_ERR_STATE, "no rollout backward has run");
  if (grad_state0 && n != c.slot_elems())
    return set_error(CN_ERR_INVALID_ARG, "grad_state0 must hold %zu values", c.slot_elems());
  if ((grad_mu || grad_sigma || grad_w) && K != c.K) return set_error(CN_ERR_INVALID_ARG, "expected K = %u", c.K);
  CN_HIP(hipSetDevice(e->device));
  const hipMemcpyKind kind = device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (grad_state0)
    CN_HIP(hipMemcpyAsync(grad_state0, c.dev.grad + (size_t)e->res.grad_slot * c.slot_elems(), n * 4, kind,
                          e->stream));
  float* outs[3] = {grad_mu, grad_sigma, grad_w};
  for (int j = 0; j < 3; ++j)
    if (outs[j]) CN_HIP(hipMemcpyAsync(outs[j], c.dev.gparam + (size_t)j * c.K, c.K * 4, kind, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  return CN_OK;
  CN_CATCH
}

// ---------------------------------------------------------------------------------------
// batch-wide parameter copies and the device-side optimiser step
// ---------------------------------------------------------------------------------------
}  // extern "C"

namespace {

// Uploads every size class's descriptor array from the host copies (one copy per class).
cn_status upload_all_descs(cn_engine* e) {
  for (const SizeClass& k : e->cls) {
    std::vector<CreatureDev> tmp(k.members.size());
    for (size_t j = 0; j < k.members.size(); ++j) tmp[j] = e->cr[k.members[j]].dev;
    CN_HIP(hipMemcpyAsync(k.d_descs, tmp.data(), tmp.size() * sizeof(CreatureDev), hipMemcpyHostToDevice, e->stream));
    CN_HIP(hipStreamSynchronize(e->stream));
  }
  return CN_OK;
}

// Adam state (zeroed moments) of every creature, allocated on first use.
cn_status ensure_adam(cn_engine* e) {
  bool grown = false;
  for (auto& c : e->cr) {
    if (c.dev.adam) continue;
    void* p = nullptr;
    const cn_status st = dalloc(e, c.allocs, &p, (6 * (size_t)c.K + 2) * 4);
    if (st != CN_OK) return st;
    c.dev.adam = (float*)p;
    grown = true;
  }
  if (!grown) return CN_OK;
  cn_status st = refresh_host_growth(e);
  if (st != CN_OK) return st;
  if ((st = upload_all_descs(e)) != CN_OK) return st;
  for (const SizeClass& k : e->cls)
    CN_HIP(launch_adam_reset(k.d_descs, (int)k.members.size(), e->stream));
  return CN_OK;
}

}  // namespace

extern "C" {

cn_status cn_batch_get_growth(cn_engine* e, float* out, size_t n) {
  CN_TRY
  clear_error();
  if (!e || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (n != e->nparams) return set_error(CN_ERR_INVALID_ARG, "expected %zu values (3K per creature)", e->nparams);
  CN_HIP(hipSetDevice(e->device));
  const cn_status st = refresh_host_growth(e);
  if (st != CN_OK) return st;
  for (const auto& c : e->cr) {
    float* o = out + c.dev.param_ofs;
    std::memcpy(o, c.dev.mu, c.K * 4);
    std::memcpy(o + c.K, c.dev.sigma, c.K * 4);
    std::memcpy(o + 2 * c.K, c.dev.w, c.K * 4);
  }
  return CN_OK;
  CN_CATCH
}

cn_status cn_batch_set_growth(cn_engine* e, const float* params, size_t n) {
  CN_TRY
  clear_error();
  if (!e || !params) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (n != e->nparams) return set_error(CN_ERR_INVALID_ARG, "expected %zu values (3K per creature)", e->nparams);
  for (size_t i = 0; i < e->cr.size(); ++i) {
    const auto& c = e->cr[i];
    const float* q = params + c.dev.param_ofs;
    const cn_status st = validate_growth(c.K, q, q + c.K, q + 2 * c.K);
    if (st != CN_OK) {
      char msg[kErrorTextSize];
      std::memcpy(msg, last_error_text(), sizeof msg);
      return set_error(st, "creature %zu: %s", i, msg);
    }
  }
  CN_HIP(hipSetDevice(e->device));
  cn_status st = refresh_host_growth(e);  // the device copies are about to be replaced anyway
  if (st != CN_OK) return st;
  for (auto& c : e->cr) {
    const float* q = params + c.dev.param_ofs;
    fill_growth(c.dev, c.K, q, q + c.K, q + 2 * c.K);
  }
  e->res.fwd_valid = false;
  return upload_all_descs(e);
  CN_CATCH
}

cn_status cn_batch_get_param_grads(cn_engine* e, float* out, size_t n) {
  CN_TRY
  clear_error();
  if (!e || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (n != e->nparams) return set_error(CN_ERR_INVALID_ARG, "expected %zu values (3K per creature)", e->nparams);
  if (!e->res.grads_valid) return set_error(CN_ERR_STATE, "no rollout backward has run");
  CN_HIP(hipSetDevice(e->device));
  if (!e->grad_gather) {
    void* p = nullptr;
    const cn_status st = dalloc(e, e->allocs, &p, e->nparams * 4);
    if (st != CN_OK) return st;
    e->grad_gather = (float*)p;
  }
  for (const SizeClass& k : e->cls)
    CN_HIP(launch_gather_grads(k.d_descs, (int)k.members.size(), e->grad_gather, e->stream));
  CN_HIP(hipMemcpyAsync(out, e->grad_gather, n * 4, hipMemcpyDeviceToHost, e->stream));
  CN_HIP(hipStreamSynchronize(e->stream));
  return CN_OK;
  CN_CATCH
}

cn_status cn_adam_reset(cn_engine* e) {
  CN_TRY
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  CN_HIP(hipSetDevice(e->device));
  const bool had = !e->cr.empty() && e->cr[0].dev.adam;
  const cn_status st = ensure_adam(e);  // a fresh allocation starts reset
  if (st != CN_OK) return st;
  if (had)
    for (const SizeClass& k : e->cls)
      CN_HIP(launch_adam_reset(k.d_descs, (int)k.members.size(), e->stream));
  return CN_OK;
  CN_CATCH
}

cn_status cn_adam_step(cn_engine* e, const cn_adam_params* p) {
  CN_TRY
  clear_error();
  if (!e || !p) return set_error(CN_ERR_INVALID_ARG, "null argument");
  const float v[8] = {p->lr, p->beta1, p->beta2, p->eps, p->mu_min, p->mu_max, p->sigma_min, p->weight_min};
  for (float x : v)
    if (!std::isfinite(x)) return set_error(CN_ERR_INVALID_ARG, "Adam parameters must be finite");
  if (!(p->lr > 0.f) || !(p->eps > 0.f)) return set_error(CN_ERR_INVALID_ARG, "lr and eps must be > 0");
  if (!(p->beta1 >= 0.f && p->beta1 < 1.f) || !(p->beta2 >= 0.f && p->beta2 < 1.f))
    return set_error(CN_ERR_INVALID_ARG, "beta1 and beta2 must be in [0, 1)");
  if (!(p->mu_min <= p->mu_max)) return set_error(CN_ERR_INVALID_ARG, "mu_min must be <= mu_max");
  if (!(p->sigma_min > 0.f)) return set_error(CN_ERR_INVALID_ARG, "sigma_min must be > 0 (sigma stays positive)");
  if (!e->res.grads_valid) return set_error(CN_ERR_STATE, "no rollout backward has run");
  CN_HIP(hipSetDevice(e->device));
  const cn_status st = ensure_adam(e);
  if (st != CN_OK) return st;
  const AdamParams ap{p->lr, p->beta1, p->beta2, p->eps, p->mu_min, p->mu_max, p->sigma_min, p->weight_min};
  for (const SizeClass& k : e->cls) CN_HIP(launch_adam(k.d_descs, (int)k.members.size(), ap, e->stream));
  e->growth_on_device = true;
  e->res.fwd_valid = false;  // later backwards must not mix in the new parameters
  return CN_OK;
  CN_CATCH
}

// ---------------------------------------------------------------------------------------
// autotuner: picks, per size class, the FFT geometry (values per thread of the forward and
// inverse row transforms and of the column transform, column tile width) and the diffusion
// tile height by timing every candidate on the engine's own buffers with HIP events, then
// the combinations of each kernel's fastest candidates as whole steps.
// ---------------------------------------------------------------------------------------
}  // extern "C"

namespace {

bool col_geom_ok(int N, int ce, int tc) {
  const int E = std::min(ce, N), T = N / E;
  return tc <= N / 2 && tc * T <= 1024 && (long)(N + N / 16) * tc * 8 <= kColLdsBudget;  // cfg.h
}

struct Geom {
  int fwd_e, inv_e, col_e, col_tc, diff_th;
  bool col_split;
  bool col_full = false;
};

// Launches one step of size class k with geometry g and events around its kernels; returns
// the four kernel times.  The step reads the current slot and writes the other one (scratch
// here: the current state is not advanced).
// Kernel times of one step of size class k with geometry g.  An untimed step is queued first
// so that the timed kernels start behind queued work, as in graph replay, and the first one's
// time does not include the launch latency of an idle queue.
cn_status time_class_step(cn_engine* e, SizeClass& k, const Geom& g, float ms[4]) {
  StepLaunch sl = class_launch(e, k, e->cur, e->cur ^ 1);
  sl.fwd_e = g.fwd_e;
  sl.inv_e = g.inv_e;
  sl.col_e = g.col_e;
  sl.col_tc = g.col_tc;
  sl.diff_th = g.diff_th;
  sl.col_split = g.col_split;
  SizeClass kg = k;
  kg.col_full = g.col_full;
  set_seg_grid(e, kg, 0, (int)k.members.size(), sl);
  CN_HIP(launch_step(k.N, sl, e->stream));
  for (int q = 0; q < 5; ++q) sl.ev[q] = e->ev[q];
  CN_HIP(launch_step(k.N, sl, e->stream));
  CN_HIP(hipEventSynchronize(e->ev[4]));
  for (int q = 0; q < 4; ++q) CN_HIP(hipEventElapsedTime(&ms[q], e->ev[q], e->ev[q + 1]));
  return CN_OK;
}

float median(std::vector<float> v) {
  std::sort(v.begin(), v.end());
  return v.empty() ? 1e30f : v[v.size() / 2];
}

cn_status tune_class(cn_engine* e, SizeClass& k, uint32_t reps) {
  const int N = k.N;
  const std::vector<int> rows = {8, 16};
  struct ColGeom {
    int e, tc;
    bool split;
    bool full;
  };
  // Full-length column geometries, and the segmented pass where some creature has segments
  // (creatures without run the default full-length geometry beside it).
  std::vector<ColGeom> cols;
  for (bool split : {false, true}) {
    if (split && k.max_k <= k.max_c) continue;  // one kernel per source: the same launch
    for (int ce : {8, 16})
      for (int tc : {2, 4, 8, 16})
        if (col_geom_ok(N, ce, tc)) cols.push_back({ce, tc, split, true});
  }
  {
    StepLaunch probe;
    SizeClass kg = k;
    kg.col_full = false;
    set_seg_grid(e, kg, 0, (int)k.members.size(), probe);
    if (probe.seg_g > 0) cols.push_back({0, 0, false, false});
  }
  const std::vector<int> diffs = {8, 16, 32};
  std::vector<std::vector<float>> tf(rows.size()), ti(rows.size()), tcol(cols.size()), td(diffs.size());
  const size_t rounds = std::max({rows.size(), cols.size(), diffs.size()});
  // one untimed pass per candidate set to load code objects and warm caches
  for (size_t r = 0; r < rounds * (reps + 1); ++r) {
    const size_t i = r % rounds;
    const ColGeom& cg = cols[i % cols.size()];
    const Geom g{rows[i % rows.size()], rows[(i + 1) % rows.size()], cg.e, cg.tc, diffs[i % diffs.size()], cg.split,
                 cg.full};
    float ms[4];
    const cn_status st = time_class_step(e, k, g, ms);
    if (st != CN_OK) return st;
    if (r < rounds) continue;
    tf[i % rows.size()].push_back(ms[0]);
    tcol[i % cols.size()].push_back(ms[1]);
    ti[(i + 1) % rows.size()].push_back(ms[2]);
    td[i % diffs.size()].push_back(ms[3]);
  }
  // Candidates of each kernel, fastest first by their own median.
  auto ranked = [](const std::vector<std::vector<float>>& t, size_t keep) {
    std::vector<size_t> idx(t.size());
    for (size_t j = 0; j < idx.size(); ++j) idx[j] = j;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return median(t[a]) < median(t[b]); });
    idx.resize(std::min(keep, idx.size()));
    return idx;
  };
  // Joint choice: the kernels of a step share the L2 and run back to back, so the per-kernel
  // medians only shortlist (3 column geometries, 2 of each other kernel); every combination of
  // the shortlists is then timed as whole steps, round robin, and the fastest median step wins.
  const std::vector<size_t> rf = ranked(tf, 2), ri = ranked(ti, 2), rc = ranked(tcol, 3), rd = ranked(td, 2);
  std::vector<Geom> combos;
  for (size_t c : rc)
    for (size_t f : rf)
      for (size_t i : ri)
        for (size_t d : rd)
          combos.push_back({rows[f], rows[i], cols[c].e, cols[c].tc, diffs[d], cols[c].split, cols[c].full});
  std::vector<std::vector<float>> tstep(combos.size());
  for (uint32_t r = 0; r < reps; ++r)
    for (size_t j = 0; j < combos.size(); ++j) {
      float ms[4];
      const cn_status st = time_class_step(e, k, combos[j], ms);
      if (st != CN_OK) return st;
      tstep[j].push_back(ms[0] + ms[1] + ms[2] + ms[3]);
    }
  const Geom& g = combos[ranked(tstep, 1)[0]];
  k.fwd_e = g.fwd_e;
  k.inv_e = g.inv_e;
  k.col_e = g.col_e;
  k.col_tc = g.col_tc;
  k.col_split = g.col_split;
  k.col_full = g.col_full;
  k.diff_th = g.diff_th;
  return CN_OK;
}

}  // namespace

extern "C" {

cn_status cn_autotune(cn_engine* e, uint32_t reps) {
  CN_TRY
  clear_error();
  if (!e) return set_error(CN_ERR_INVALID_ARG, "null engine");
  if (reps == 0 || reps > 1000) return set_error(CN_ERR_INVALID_ARG, "reps must be in [1, 1000]");
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipStreamSynchronize(e->stream));
  for (SizeClass& k : e->cls) {
    const cn_status st = tune_class(e, k, reps);
    if (st != CN_OK) return st;
  }
  drop_graphs(e);        // graphs bake the launch geometry in
  e->have_prev = false;  // the other slot was used as scratch
  e->res.fwd_valid = false;
  return CN_OK;
  CN_CATCH
}

uint32_t cn_engine_size_classes(const cn_engine* e) { return e ? (uint32_t)e->cls.size() : 0; }

cn_status cn_engine_get_plan(const cn_engine* e, uint32_t size_class, cn_plan* out) {
  CN_TRY
  clear_error();
  if (!e || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (size_class >= e->cls.size()) return set_error(CN_ERR_INVALID_ARG, "size class %u out of range", size_class);
  const SizeClass& k = e->cls[size_class];
  out->n = (uint32_t)k.N;
  out->fwd_e = (uint32_t)k.fwd_e;
  out->inv_e = (uint32_t)k.inv_e;
  out->col_e = (uint32_t)k.col_e;
  out->col_tc = (uint32_t)k.col_tc;
  out->diff_th = (uint32_t)k.diff_th;
  out->col_split = k.col_split ? 1u : 0u;
  out->col_full = k.col_full ? 1u : 0u;
  return CN_OK;
  CN_CATCH
}

cn_status cn_engine_set_plan(cn_engine* e, const cn_plan* plan) {
  CN_TRY
  clear_error();
  if (!e || !plan) return set_error(CN_ERR_INVALID_ARG, "null argument");
  SizeClass* k = nullptr;
  for (SizeClass& c : e->cls)
    if (c.N == (int)plan->n) k = &c;
  if (!k) return set_error(CN_ERR_INVALID_ARG, "no creature with grid size %u", plan->n);
  auto row_ok = [](uint32_t v) { return v == 0 || v == 8 || v == 16; };
  if (!row_ok(plan->fwd_e) || !row_ok(plan->inv_e) || !row_ok(plan->col_e))
    return set_error(CN_ERR_INVALID_ARG, "values per thread must be 0 (default), 8 or 16");
  if (plan->diff_th != 0 && plan->diff_th != 8 && plan->diff_th != 16 && plan->diff_th != 32)
    return set_error(CN_ERR_INVALID_ARG, "diffusion tile height must be 0, 8, 16 or 32");
  if (plan->col_split > 1) return set_error(CN_ERR_INVALID_ARG, "col_split must be 0 or 1");
  if (plan->col_full > 1) return set_error(CN_ERR_INVALID_ARG, "col_full must be 0 or 1");
  if (plan->col_tc != 0 || plan->col_e != 0) {
    const int ce = plan->col_e ? (int)plan->col_e : 16;  // the launcher's defaults (DefaultGeom)
    const int tc = plan->col_tc ? (int)plan->col_tc : default_col_tc(k->N);
    if (!col_geom_ok(k->N, ce, tc))
      return set_error(CN_ERR_INVALID_ARG, "column geometry (E=%d, TC=%d) does not fit N=%d", ce, tc, k->N);
  }
  k->fwd_e = (int)plan->fwd_e;
  k->inv_e = (int)plan->inv_e;
  k->col_e = (int)plan->col_e;
  k->col_tc = (int)plan->col_tc;
  k->diff_th = (int)plan->diff_th;
  k->col_split = plan->col_split != 0;
  k->col_full = plan->col_full != 0;
  CN_HIP(hipSetDevice(e->device));
  CN_HIP(hipStreamSynchronize(e->stream));
  drop_graphs(e);
  return CN_OK;
  CN_CATCH
}

}  // extern "C"
