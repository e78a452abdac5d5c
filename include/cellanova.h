/* cellanova.h - C ABI of the MI355X DiffusionLenia step engine.
 *
 * Drop-in boundary for the reference orchestrator `GpuMaceLenia` (ml-rs/orchestrator/mod.rs)
 * and its loader (ml-rs/config.rs).  Each entry point cites the reference interface it
 * replaces.  Conventions:
 *   - every call returns cn_status (CN_OK == 0); nothing aborts across the boundary and the
 *     last error text of the calling thread is available from cn_last_error();
 *   - arguments are validated on the host before any HIP call is made;
 *   - host arrays are caller-owned and their lengths are passed explicitly; device memory is
 *     owned by the engine; one HIP stream per engine;
 *   - a handle is used by one thread at a time (many handles may coexist);
 *   - cn_iterate / cn_run_steps are asynchronous (like `iterate`, mod.rs:385); downloads,
 *     cn_sync, cn_timed_iterate and cn_mass_fp64 synchronise.
 */
#ifndef CELLANOVA_H
#define CELLANOVA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cn_status {
  CN_OK = 0,
  CN_ERR_INVALID_ARG = 1,   /* bad pointer, size, index, parameter value */
  CN_ERR_UNSUPPORTED = 2,   /* valid but unsupported configuration (e.g. N not a power of two) */
  CN_ERR_HIP = 3,           /* a HIP runtime call failed */
  CN_ERR_IO = 4,            /* file missing / unreadable / truncated */
  CN_ERR_PARSE = 5,         /* malformed JSON */
  CN_ERR_OUT_OF_MEMORY = 6, /* device or host allocation failed */
  CN_ERR_STATE = 7          /* call not valid in the current engine state */
} cn_status;

typedef struct cn_engine cn_engine;

/* Engine construction parameters.  Replaces the argument list of GpuMaceLenia::new
 * (ml-rs/orchestrator/mod.rs:79-90): shape -> n (square grid, power of two, 16..2048),
 * num_channels, num_kernels, c0 -> src (input channel of kernel k), c1 -> dst (output
 * channel of kernel k), mu, sigma, weights, temp.  C <= 16 (compute_1024.wgsl:307), K <= 256. */
typedef struct cn_engine_desc {
  uint32_t n;
  uint32_t num_channels;
  uint32_t num_kernels;
  const uint32_t* src;   /* [K] */
  const uint32_t* dst;   /* [K] */
  const float* mu;       /* [K] */
  const float* sigma;    /* [K], > 0 */
  const float* weights;  /* [K] */
  float temp;
  int device;            /* HIP device ordinal */
} cn_engine_desc;

/* GpuMaceLenia::new (mod.rs:79-198).  All state starts at zero. */
cn_status cn_engine_create(const cn_engine_desc* desc, cn_engine** out);
/* Drop of GpuMaceLenia (buffers released). */
cn_status cn_engine_destroy(cn_engine* e);

/* GpuMaceLenia::set_kernel (mod.rs:207-214): full N*N complex64 spectrum of kernel k as
 * interleaved (re, im) f32, row-major, unnormalised FFT of the origin-centred spatial kernel.
 * n_complex must equal N*N (the reference asserts this, mod.rs:208). */
cn_status cn_set_kernel(cn_engine* e, uint32_t k, const float* interleaved, size_t n_complex);

/* GpuMaceLenia::upload_channel (mod.rs:217-226): f64 values cast to f32; n must equal N*N. */
cn_status cn_upload_channel_f64(cn_engine* e, uint32_t channel, const double* data, size_t n);
/* Same with f32 input (no cast). */
cn_status cn_upload_channel_f32(cn_engine* e, uint32_t channel, const float* data, size_t n);

/* GpuMaceLenia::iterate (mod.rs:348-386): one step, asynchronous. */
cn_status cn_iterate(cn_engine* e);
/* GpuMaceLenia::run_steps (mod.rs:229-233): n steps, asynchronous (HIP-graph replay). */
cn_status cn_run_steps(cn_engine* e, uint64_t n);
/* Wait for all queued work of the engine (device.poll(Wait) in the reference). */
cn_status cn_sync(cn_engine* e);

/* GpuMaceLenia::iterate_with_encoder (mod.rs:435-466), which records one step into the
 * caller's command encoder: here one step is enqueued on the caller's HIP stream `stream`
 * (a hipStream_t; NULL is the null stream) instead of the engine's.  The caller's stream first
 * waits for the work already queued on the engine's stream, and the engine's stream waits for
 * the step, so both orders hold without a host wait.  Asynchronous. */
cn_status cn_iterate_on_stream(cn_engine* e, void* stream);
/* New: the engine's own HIP stream (a hipStream_t), for callers that enqueue their work behind
 * or ahead of the engine's with events.  Valid until cn_engine_destroy. */
cn_status cn_engine_stream(const cn_engine* e, void** stream);
/* GpuMaceLenia::channel_buffer (mod.rs:202-204): device address of creature i's current state,
 * C*N*N f32, channel-major (*n receives C*N*N if n is not NULL).  The engine ping-pongs between
 * two state buffers, so the address is valid until the next step or rollout call; work on
 * other streams must be ordered with the engine's stream (cn_engine_stream).  Writing through
 * it replaces the state like cn_creature_state_from_device, except that the engine is not told:
 * cn_download_conv_real and cn_rollout_backward then refer to the state before the write. */
cn_status cn_creature_state_ptr(cn_engine* e, uint32_t creature, float** d_state, size_t* n);

/* GpuMaceLenia::timed_iterate (mod.rs:469-540): one step with per-phase device times in ms:
 * ms[0] convolution (row + fused column passes), ms[1] growth (inverse row pass fused with
 * the growth/affinity epilogue), ms[2] diffusion, ms[3] copy (0: ping-pong pointer swap). */
cn_status cn_timed_iterate(cn_engine* e, double ms[4]);

/* New (extends timed_iterate): 4 * steps directly launched steps queued back to back on the
 * engine's stream, step s with HIP events around pass s % 4 (first size class) only; returns the
 * average device time per launch in ms: kernel_ms[0] row forward (R2C), [1] column pass, [2]
 * inverse row + growth, [3] diffusion.  The state advances 4 * steps steps.  n must be 4,
 * 1 <= steps <= 10000.  Synchronises. */
cn_status cn_profile_kernels(cn_engine* e, uint64_t steps, double* kernel_ms, size_t n);

/* GpuMaceLenia::download_all_channels (mod.rs:236-263): C*N*N f32, channel-major. */
cn_status cn_download_channels(cn_engine* e, float* out, size_t n);
/* GpuMaceLenia::download_conv_buffer (mod.rs:267-299): K*N*N f32, the real part of the
 * UNNORMALISED inverse transform (N^2 times the convolution) computed from the state the
 * most recent step started from; zeros before the first step. */
cn_status cn_download_conv_real(cn_engine* e, float* out, size_t n);

/* GpuMaceLenia::iterate_and_render (mod.rs:390-432): one step whose diffusion pass also packs
 * channels 0..2 of the new state into a 0x00RRGGBB frame, pixel = (r << 16) | (g << 8) | b with
 * each byte u32(sqrt(clamp(1.5 v, 0, 1)) * 255) (render.rs:145-153, compute_1024.wgsl:429-443;
 * IEEE f32 sqrt, truncation; channels >= C read as 0).  Every creature of a batch is rendered.
 * Asynchronous like cn_iterate. */
cn_status cn_iterate_and_render(cn_engine* e);

/* GpuMaceLenia::download_render (mod.rs:302-335): the last rendered frame as RGB24 bytes, row
 * major, n = 3 N^2.  All zeros before the first cn_iterate_and_render.  Synchronises. */
cn_status cn_download_render(cn_engine* e, uint8_t* rgb, size_t n);
cn_status cn_creature_download_render(cn_engine* e, uint32_t i, uint8_t* rgb, size_t n);

/* New: per-channel total mass of the current state, FP64 Neumaier-compensated sums on the
 * device (out[c], c < C).  Synchronises. */
cn_status cn_mass_fp64(cn_engine* e, double* out, size_t n_channels);

/* New, with no counterpart in the reference: the FP64 mass of a caller's device state for
 * callers that cannot allocate (code under HIP graph capture, serving loops that forbid
 * allocation).  d_state: num_channels planes of n*n f32 in device memory (channel-major, like
 * cn_creature_state_ptr's); d_out: num_channels doubles in device memory, the same
 * Neumaier-compensated sums as cn_creature_mass_fp64 gives for that state, bit for bit.
 * Enqueued on `stream` (a hipStream_t; NULL is the null stream) of the calling thread's current
 * device; asynchronous.  Makes no allocation: the block partials live in a fixed device array
 * of 16 KiB per device, so num_channels is at most 16 (n at most 65536), and calls that overlap
 * on one device race on that array. */
cn_status cn_mass_fp64_noalloc(const float* d_state, uint32_t n, uint32_t num_channels, double* d_out,
                               void* stream);

/* New: engine facts. */
typedef struct cn_engine_info {
  uint32_t n, num_channels, num_kernels;
  uint64_t steps_done;
  uint64_t device_bytes;   /* device memory held by the engine */
  uint32_t compact_kernels; /* kernels of creature 0 whose spectrum is real and even in ky and
                               is read from the compact real copy (a quarter of the bytes) */
  uint32_t segments;        /* creature 0's segmented column pass: segments per column (0: its
                               column transforms are full-length) */
  uint32_t segment_radius;  /* the largest column-profile support radius of its kernels: the
                               segments overlap by twice this many rows */
  uint32_t segment_length;  /* rows per segment (the segment transform length) */
} cn_engine_info;
cn_status cn_engine_get_info(const cn_engine* e, cn_engine_info* out);

/* Text of the calling thread's last error ("" if none).  The pointer stays valid until the
 * next cn_* call on the same thread. */
const char* cn_last_error(void);

/* ------------------------------------------------------------------------------------ */
/* Batches (no reference counterpart: many creatures of mixed N, C, K in one engine)      */
/* ------------------------------------------------------------------------------------ */

/* An engine holding `count` creatures (descs[i].device is ignored; `device` is used).  Every
 * step / rollout call advances all creatures together; creatures of equal N share one launch
 * per kernel (one descriptor table per size class) and all classes are captured in one HIP
 * graph.  The single-creature calls above act on creature 0.  At most 65535 creatures
 * (CN_ERR_UNSUPPORTED beyond). */
cn_status cn_batch_create(const cn_engine_desc* descs, uint32_t count, int device, cn_engine** out);
/* Per-creature forms of set_kernel / upload_channel / download_all_channels / mass. */
cn_status cn_creature_set_kernel(cn_engine* e, uint32_t creature, uint32_t k, const float* interleaved,
                                 size_t n_complex);
cn_status cn_creature_upload_channel_f32(cn_engine* e, uint32_t creature, uint32_t channel, const float* data,
                                         size_t n);
cn_status cn_creature_download_channels(cn_engine* e, uint32_t creature, float* out, size_t n);
cn_status cn_creature_mass_fp64(cn_engine* e, uint32_t creature, double* out, size_t n_channels);
/* Replace the growth parameters (mu, sigma > 0, weights; K each) of one creature. */
cn_status cn_creature_set_growth(cn_engine* e, uint32_t creature, const float* mu, const float* sigma,
                                 const float* weights, size_t K);
/* Copy the current state of a creature from / to caller-owned DEVICE memory (C*N*N f32).
 * Both synchronise the engine stream; the caller orders its own streams before the call. */
cn_status cn_creature_state_from_device(cn_engine* e, uint32_t creature, const float* d_src, size_t n);
cn_status cn_creature_state_to_device(cn_engine* e, uint32_t creature, float* d_dst, size_t n);

/* ------------------------------------------------------------------------------------ */
/* Differentiable rollouts (no reference counterpart; the reference differentiates its    */
/* PyTorch model, train/diff_lenia_org.py:27-53, with autograd)                           */
/* ------------------------------------------------------------------------------------ */

/* Loss target of a creature (C*N*N f32, host memory) for cn_rollout_grad. */
cn_status cn_creature_set_target(cn_engine* e, uint32_t creature, const float* target, size_t n);
/* Rollout of `steps` steps of every creature from its current state, keeping a checkpoint
 * every `ckpt_every` steps, then the MSE loss mean((x_T - target)^2) and the checkpointed
 * adjoint (recomputing each segment from its checkpoint): dL/dstate0 and dL/d(mu, sigma,
 * weights) summed over the steps.  The current state is not modified.  Asynchronous; one HIP
 * graph per (steps, ckpt_every) is captured on first use and replayed afterwards. */
cn_status cn_rollout_grad(cn_engine* e, uint32_t steps, uint32_t ckpt_every);
/* Forward half only (keeps the checkpoints for cn_rollout_backward).  Asynchronous. */
cn_status cn_rollout_forward(cn_engine* e, uint32_t steps, uint32_t ckpt_every);
/* Final state of the last rollout forward (host memory, or device memory if device_pointer). */
cn_status cn_creature_rollout_final(cn_engine* e, uint32_t creature, float* out, size_t n, int device_pointer);
/* Adjoint of the last cn_rollout_forward given dL/dx_T per creature (DEVICE pointers, one per
 * creature, C*N*N f32 each).  Synchronises. */
cn_status cn_rollout_backward(cn_engine* e, const float* const* d_grad_final, uint32_t count);
/* MSE loss of the last cn_rollout_grad (synchronises). */
cn_status cn_creature_get_loss(cn_engine* e, uint32_t creature, double* loss);
/* Gradients of the last backward: dL/dstate0 (C*N*N, n must match; may be NULL) and
 * dL/dmu, dL/dsigma, dL/dweights (K each; may be NULL).  Host pointers, or device pointers if
 * device_pointers != 0.  Synchronises. */
cn_status cn_creature_get_grads(cn_engine* e, uint32_t creature, float* grad_state0, size_t n, float* grad_mu,
                                float* grad_sigma, float* grad_w, size_t K, int device_pointers);

/* ------------------------------------------------------------------------------------ */
/* Creature search on the device (no reference counterpart: the reference runs one        */
/* creature at a time through torch.optim.Adam, train/diff_lenia_org.py, and clamps the    */
/* parameters with LeniaParams._sanitize, train/utils/leniaparams.py:351-360)            */
/* ------------------------------------------------------------------------------------ */

/* Batch-wide parameter arrays: creature i's 3K values [mu K | sigma K | weights K] follow
 * those of creature i-1; n is the sum over creatures of 3K.  get/set copy all creatures at
 * once (one synchronisation); get_param_grads returns dL/d(mu, sigma, w) of the last backward. */
cn_status cn_batch_get_growth(cn_engine* e, float* out, size_t n);
cn_status cn_batch_set_growth(cn_engine* e, const float* params, size_t n);
cn_status cn_batch_get_param_grads(cn_engine* e, float* out, size_t n);

/* Adam on every creature's (mu, sigma, w), on the device, from the gradients of the last
 * rollout backward: m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2, step t counted per call,
 * p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps) in FP32, then mu is clamped to
 * [mu_min, mu_max], sigma to >= sigma_min (> 0) and the weights to >= weight_min.  The new
 * parameters are written where the step kernels read them: a loop of cn_rollout_grad and
 * cn_adam_step runs with no host round trip.  Asynchronous.  The moments start at zero on
 * the first call and after cn_adam_reset. */
typedef struct cn_adam_params {
  float lr, beta1, beta2, eps;
  float mu_min, mu_max, sigma_min, weight_min;
} cn_adam_params;
cn_status cn_adam_step(cn_engine* e, const cn_adam_params* p);
cn_status cn_adam_reset(cn_engine* e);

/* ------------------------------------------------------------------------------------ */
/* Launch plans and the autotuner (no reference counterpart: the reference hard-codes one */
/* shader per grid size, mod.rs:30-39)                                                    */
/* ------------------------------------------------------------------------------------ */

/* Geometry of one grid-size class: values per thread (8 or 16) of the forward row (R2C),
 * inverse row (C2R) and column FFTs, column tile width of the column passes (2..16), and the
 * diffusion tile height (8, 16, 32).  0 selects the built-in default.  col_split = 1 runs the
 * forward column pass as one block per (tile, kernel) instead of per (tile, source channel)
 * (more blocks for small grids; each block redoes its source's forward column transform). */
typedef struct cn_plan {
  uint32_t n;
  uint32_t fwd_e, inv_e, col_e, col_tc, diff_th;
  uint32_t col_split;
  /* 1: full-length column transforms for every creature of the class; 0: the segmented column
   * pass for creatures whose kernels have real, even spectra of compact column support (see
   * cn_engine_info's segments), full-length transforms for the others. */
  uint32_t col_full;
} cn_plan;

/* Times every candidate geometry of every size class on the device (reps timed launches per
 * candidate, median), shortlists the fastest per kernel (3 column geometries, 2 of each other
 * kernel), times every combination of the shortlists as whole steps (reps each, median) and
 * keeps the fastest step.  Synchronises; the current state is unchanged.  Captured graphs are
 * rebuilt on next use. */
cn_status cn_autotune(cn_engine* e, uint32_t reps);
uint32_t cn_engine_size_classes(const cn_engine* e);
cn_status cn_engine_get_plan(const cn_engine* e, uint32_t size_class, cn_plan* out);
/* Applies a plan to the size class with grid size plan->n (validated; graphs rebuilt). */
cn_status cn_engine_set_plan(cn_engine* e, const cn_plan* plan);

/* ------------------------------------------------------------------------------------ */
/* Loader (ml-rs/config.rs)                                                              */
/* ------------------------------------------------------------------------------------ */

/* SimConfig (config.rs:14-25) after load_config's c0 <-> c1 swap (config.rs:40). */
typedef struct cn_config {
  uint32_t seed_size;
  uint32_t num_channels;
  uint32_t num_kernels;
  double* seed_channels;   /* [num_channels][seed_size*seed_size] */
  size_t seed_len;         /* elements per channel actually present (must be seed_size^2) */
  uint32_t* c0;            /* [K] input channel (= JSON c1) */
  uint32_t* c1;            /* [K] output channel (= JSON c0) */
  float* growth_mu;        /* [K] */
  float* growth_sigma;     /* [K] */
  float* growth_weights;   /* [K] */
} cn_config;

/* load_config (config.rs:32-47): parse the JSON file and swap c0/c1. */
cn_status cn_load_config(const char* path, cn_config** out);
void cn_free_config(cn_config* cfg);

/* load_kernels (config.rs:56-125): read K spectra of S*S complex64 (S inferred from the
 * file size) and, if S != target, zero-pad them spectrally to target*target (fftshift,
 * centre pad, ifftshift).  out receives K*target*target interleaved (re, im) pairs;
 * n_floats must equal 2*K*target*target.  A file whose size is not K*S*S*8 bytes is
 * rejected (the reference would silently mis-size it, config.rs:62-65), as is target < S. */
cn_status cn_load_kernels(const char* path, uint32_t num_kernels, uint32_t target, float* out,
                          size_t n_floats);

/* Writers for the two input files (train/generate_kernel_json.py:171-213).
 * cn_write_kernels: K spectra of n*n complex64 as little-endian (re, im) f32 pairs, kernel-major,
 * row-major (the .bin format read by load_kernels, config.rs:49-55); kernels_fft holds
 * 2*K*n*n floats.
 * cn_write_config: the creature JSON (seed_size, num_channels, num_kernels, seed_channels,
 * c0, c1, growth_mu, growth_sigma, growth_weights).  cfg is in the loader's convention
 * (c0 = input channel); the file gets the JSON convention (JSON c0 = cfg->c1, JSON c1 = cfg->c0)
 * so that cn_load_config reads back cfg exactly (seeds are written with 17 significant digits). */
cn_status cn_write_kernels(const char* path, const float* kernels_fft, uint32_t num_kernels, uint32_t n);
cn_status cn_write_config(const char* path, const cn_config* cfg);

/* ------------------------------------------------------------------------------------ */
/* Synthetic creatures (no reference counterpart: the benchmark distributions)           */
/* ------------------------------------------------------------------------------------ */

/* Lenia ring-kernel creature.  Kernel k: radius R, B_k ~ U{1,2,3} rings with peaks
 * beta ~ U[0.3,1], core exp(4 - 1/(f(1-f))), normalised to sum 1 in FP64, origin at (0,0),
 * FP64 FFT rounded to complex64.  mu ~ U[0.05,0.15], sigma ~ U[0.03,0.06], weights
 * Dirichlet(1) over the kernels sharing a destination.  Seed: U[0,1) on the central
 * patch x patch square of every channel, zero elsewhere.  All draws come from a
 * counter-based generator keyed by (seed, creature, stream, index). */
typedef struct cn_creature_params {
  uint32_t n, num_channels, num_kernels;
  const uint32_t* src;      /* [K] */
  const uint32_t* dst;      /* [K] */
  const uint32_t* radius;   /* [K] kernel radii R (cells) */
  uint32_t patch;           /* seed patch side (clamped to n) */
  uint64_t seed;
  uint64_t creature;
} cn_creature_params;

/* Outputs (any may be NULL): kernels_fft [K][n*n] interleaved complex64 (2*K*n*n floats),
 * spatial [K][n*n] f64 kernels, state [C][n*n] f32, mu/sigma/weights [K]. */
cn_status cn_generate_creature(const cn_creature_params* p, float* kernels_fft, double* spatial,
                               float* state, float* mu, float* sigma, float* weights);

/* The creature of the reference's CPU example (ml-rs/example.rs:310-491: generate_kernels_fft,
 * ifftshift_2d, generate_seed, generate_growth_params), at grid size n (the example fixes
 * n = 1024; C = 3 and K = 9 always).  Kernel k: a Gaussian ring of width w_k at half the scaled
 * radius d / (10 r_k), under a sigmoid envelope ending at the scaled radius 1, with the example's
 * r and w tables; normalised to sum 1 and moved so its centre sits at the origin.  Outputs (any
 * may be NULL): kernels_fft [9][n*n] interleaved complex64 (FP64 transform, rounded), spatial
 * [9][n*n] f64 (the origin-centred kernels whose transform that is), state [3][n*n] f32 (three
 * Gaussian blobs), mu/sigma/weights [9] (the example's f32 values), src/dst [9] (k % 3, k / 3:
 * the example's c0 = input, c1 = output).  Kernels and seed are evaluated in FP64 (the example
 * uses f32; the two agree to f32 rounding). */
cn_status cn_generate_example_creature(uint32_t n, float* kernels_fft, double* spatial, float* state, float* mu,
                                       float* sigma, float* weights, uint32_t* src, uint32_t* dst);

/* New (the reference generates spectra on the host, train/generate_kernel_json.py:106-145,
 * and uploads them with set_kernel, mod.rs:207): generate creature i of an engine in place on
 * the device from the distribution above.  Kernels are evaluated and normalised in FP64 on the
 * grid and transformed by the engine's f32 FFT straight into its spectrum storage; the seed
 * state (current state) and the growth parameters are bit-identical to cn_generate_creature.
 * p->n, num_channels, num_kernels, src and dst must equal the creature's.  Synchronises. */
cn_status cn_creature_generate(cn_engine* e, uint32_t i, const cn_creature_params* p);

/* New: the kernel spectra creature i steps with, as full N*N complex64 (interleaved re, im; out
 * holds 2*K*N*N floats), unscaled like set_kernel's input.  For spectra given to set_kernel this
 * is their Hermitian part (the part a real convolution uses).  Synchronises. */
cn_status cn_creature_download_kernels(cn_engine* e, uint32_t i, float* out, size_t n_floats);

/* ------------------------------------------------------------------------------------ */
/* Audio consumer of the rendered frames (ml-rs/audio.rs, used by main.rs:219-360)         */
/* ------------------------------------------------------------------------------------ */

/* AudioSynth::extract_features (audio.rs:181-209) on the device: per RGB channel, the mean and
 * the variance (max(mean of squares - mean^2, 0), f32) of byte / 255 over the last frame that
 * cn_iterate_and_render packed for creature i, from exact integer byte sums (the reference sums
 * byte / 255 in f64 over the downloaded RGB24 bytes).  All zeros before the first rendered
 * frame.  Synchronises. */
cn_status cn_creature_render_features(cn_engine* e, uint32_t creature, float means[3], float variances[3]);
cn_status cn_render_features(cn_engine* e, float means[3], float variances[3]);
/* The same features of host RGB24 bytes (3 * grid^2 bytes, row major), as extract_features
 * computes them from `render_rgb` and `grid_size`. */
cn_status cn_audio_extract_features(const uint8_t* rgb, uint32_t grid, float means[3], float variances[3]);

/* AudioSynth (audio.rs:104-160): the sonar-ping synthesiser, one per video. */
typedef struct cn_audio cn_audio;
cn_status cn_audio_create(uint32_t sample_rate, cn_audio** out); /* AudioSynth::new */
cn_status cn_audio_destroy(cn_audio* a);
/* AudioSynth::generate_frame (audio.rs:310-396): triggers pings from the features of one video
 * frame, then synthesises num_samples stereo samples into out (n = 2 * num_samples int16,
 * interleaved left, right). */
cn_status cn_audio_generate_frame(cn_audio* a, const float means[3], const float variances[3],
                                  uint32_t num_samples, int16_t* out, size_t n);
cn_status cn_audio_activity(const cn_audio* a, float* out); /* AudioSynth::activity_level */
/* write_wav (audio.rs:404-441): 16-bit PCM RIFF/WAVE file of `count` interleaved samples. */
cn_status cn_write_wav(const char* path, const int16_t* samples, size_t count, uint32_t sample_rate,
                       uint16_t channels);

/* ------------------------------------------------------------------------------------ */
/* Test hooks                                                                            */
/* ------------------------------------------------------------------------------------ */

/* Make the next `count` HIP calls issued through the engine's checked wrapper report failure
 * without calling HIP (exercises error paths on machines without a GPU). */
void cn_debug_fail_next_hip(int count);

/* Let `skip` checked HIP calls / device allocations through, then make the next one report
 * failure (once) without calling HIP; a negative value cancels.  Walking `skip` over 0, 1, 2, ...
 * fails each step of a call in turn (error paths of the allocation and capture sequences). */
void cn_debug_fail_hip_at(int skip);

#ifdef __cplusplus
}
#endif
#endif /* CELLANOVA_H */
