"""Python mirror of the reference orchestrator `GpuMaceLenia` (ml-rs/orchestrator/mod.rs:45-541)
over the C ABI, plus the batch and rollout interfaces.  Same method set and argument meaning
as the reference; errors raise CellaNovaError instead of panicking (mod.rs:94,208,218)."""
import ctypes

import numpy as np

from ._lib import AdamParams, EngineDesc, EngineInfo, Plan, check, lib


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _desc(n, C, K, src, dst, mu, sigma, w, temp, device, keep):
    arrs = [_u32(src), _u32(dst), _f32(mu), _f32(sigma), _f32(w)]
    for a in arrs:
        if a.shape != (K,):
            raise ValueError(f"per-kernel arrays must have {K} entries")
    keep.extend(arrs)
    P = ctypes.POINTER
    return EngineDesc(int(n), int(C), int(K), arrs[0].ctypes.data_as(P(ctypes.c_uint32)),
                      arrs[1].ctypes.data_as(P(ctypes.c_uint32)), arrs[2].ctypes.data_as(P(ctypes.c_float)),
                      arrs[3].ctypes.data_as(P(ctypes.c_float)), arrs[4].ctypes.data_as(P(ctypes.c_float)),
                      float(temp), int(device))


def _dev_f32(t, numel, what):
    """Device address of `t`, which must be a contiguous float32 device tensor of `numel`
    elements (the C ABI copies numel * 4 raw bytes)."""
    if not hasattr(t, "data_ptr") or not t.is_cuda:
        raise TypeError(f"{what} must be a device tensor")
    if str(t.dtype) != "torch.float32" or not t.is_contiguous():
        raise TypeError(f"{what} must be a contiguous float32 tensor, got {t.dtype}")
    if t.numel() != numel:
        raise ValueError(f"{what} has {t.numel()} elements, expected {numel}")
    return t.data_ptr()


def _host_f32(a, numel, what):
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or not a.flags.c_contiguous or a.size != numel:
        raise TypeError(f"{what} must be a contiguous float32 numpy array of {numel} elements")
    return a.ctypes.data


class _Engine:
    """Creature-indexed operations shared by GpuMaceLenia (one creature) and CreatureBatch."""

    _h = None

    def _init(self, descs, device):
        self._shapes = [(int(d.num_channels), int(d.n), int(d.num_kernels)) for d in descs]
        self._maps = [(np.ctypeslib.as_array(d.src, shape=(d.num_kernels,)).copy(),
                       np.ctypeslib.as_array(d.dst, shape=(d.num_kernels,)).copy()) for d in descs]
        h = ctypes.c_void_p()
        if len(descs) == 1:
            check(lib().cn_engine_create(ctypes.byref(descs[0]), ctypes.byref(h)))
        else:
            arr = (EngineDesc * len(descs))(*descs)
            check(lib().cn_batch_create(arr, len(descs), int(device), ctypes.byref(h)))
        self._h = h

    # --- lifecycle -------------------------------------------------------------------------
    def close(self):
        if self._h is not None and self._h.value:
            check(lib().cn_engine_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _elems(self, i):
        C, n, _ = self._shapes[i]
        return C * n * n

    # --- per creature ------------------------------------------------------------------------
    def creature_set_kernel(self, i, k, kernel_fft):
        a = np.asarray(kernel_fft)
        if np.iscomplexobj(a):
            a = np.ascontiguousarray(a.astype(np.complex64)).view(np.float32)
        else:
            a = _f32(a)
        a = a.reshape(-1)
        check(lib().cn_creature_set_kernel(self._h, int(i), int(k), a.ctypes.data, a.size // 2))

    def creature_upload_channel(self, i, c, data):
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float32)).reshape(-1)
        check(lib().cn_creature_upload_channel_f32(self._h, int(i), int(c), a.ctypes.data, a.size))

    def creature_download(self, i):
        out = np.empty(self._elems(i), dtype=np.float32)
        check(lib().cn_creature_download_channels(self._h, int(i), out.ctypes.data, out.size))
        return out

    def creature_mass(self, i):
        out = np.empty(self._shapes[i][0], dtype=np.float64)
        check(lib().cn_creature_mass_fp64(self._h, int(i), out.ctypes.data, out.size))
        return out

    def creature_set_growth(self, i, mu, sigma, w):
        m, s, ww = _f32(mu), _f32(sigma), _f32(w)
        check(lib().cn_creature_set_growth(self._h, int(i), m.ctypes.data, s.ctypes.data, ww.ctypes.data, m.size))

    def creature_state_from_device(self, i, tensor):
        """Copy a contiguous f32 device tensor (C, N, N) into the creature's current state."""
        n = self._elems(i)
        check(lib().cn_creature_state_from_device(self._h, int(i), _dev_f32(tensor, n, "tensor"), n))

    def creature_state_to_device(self, i, tensor):
        n = self._elems(i)
        check(lib().cn_creature_state_to_device(self._h, int(i), _dev_f32(tensor, n, "tensor"), n))

    def creature_set_target(self, i, target):
        a = _f32(target).reshape(-1)
        check(lib().cn_creature_set_target(self._h, int(i), a.ctypes.data, a.size))

    def creature_rollout_final(self, i, out=None):
        """Final state of the last rollout: into `out` (device tensor) or a new numpy array."""
        if out is None:
            arr = np.empty(self._elems(i), dtype=np.float32)
            check(lib().cn_creature_rollout_final(self._h, int(i), arr.ctypes.data, arr.size, 0))
            return arr
        n = self._elems(i)
        check(lib().cn_creature_rollout_final(self._h, int(i), _dev_f32(out, n, "out"), n, 1))
        return out

    def creature_loss(self, i):
        v = ctypes.c_double()
        check(lib().cn_creature_get_loss(self._h, int(i), ctypes.byref(v)))
        return v.value

    def creature_grads(self, i, state=None, mu=None, sigma=None, w=None, device=False, with_state=True):
        """Gradients of the last backward.  Host: returns numpy arrays (dL/dstate0 only if
        with_state).  device=True: fills the given device tensors."""
        C, n, K = self._shapes[i]
        ptr = _dev_f32 if device else _host_f32
        if not device:
            state = np.empty(C * n * n, np.float32) if with_state else None
            mu, sigma, w = (np.empty(K, np.float32) for _ in range(3))
        check(lib().cn_creature_get_grads(
            self._h, int(i), ptr(state, C * n * n, "state") if state is not None else None, C * n * n,
            ptr(mu, K, "mu") if mu is not None else None, ptr(sigma, K, "sigma") if sigma is not None else None,
            ptr(w, K, "w") if w is not None else None, K, 1 if device else 0))
        return {"state0": state, "mu": mu, "sigma": sigma, "weights": w}

    def creature_generate(self, i, radius, patch=128, seed=0, creature=0, src=None, dst=None):
        """Generate creature i in place on the device (kernel spectra, growth parameters, seed
        state) from the synthetic distribution; src/dst default to the creature's own maps."""
        from .creatures import creature_params
        C, n, K = self._shapes[i]
        if src is None:
            src, dst = self._maps[i]
        p, keep = creature_params(n, C, K, src, dst, radius, patch, seed, creature)
        check(lib().cn_creature_generate(self._h, int(i), ctypes.byref(p)))
        del keep

    def creature_download_kernels(self, i):
        """The spectra creature i steps with: (K, N*N) complex64, unscaled (set_kernel's format)."""
        C, n, K = self._shapes[i]
        out = np.empty(2 * K * n * n, dtype=np.float32)
        check(lib().cn_creature_download_kernels(self._h, int(i), out.ctypes.data, out.size))
        return out.view(np.complex64).reshape(K, n * n)

    def creature_download_render(self, i):
        """RGB24 frame of the last iterate_and_render, (N, N, 3) uint8."""
        n = self._shapes[i][1]
        out = np.empty((n, n, 3), dtype=np.uint8)
        check(lib().cn_creature_download_render(self._h, int(i), out.ctypes.data, out.size))
        return out

    def creature_render_features(self, i):
        """(means, variances), float32[3] each: the audio features of creature i's last rendered
        frame, reduced on the device (audio.rs:181-209)."""
        means = np.empty(3, dtype=np.float32)
        variances = np.empty(3, dtype=np.float32)
        check(lib().cn_creature_render_features(self._h, int(i), means.ctypes.data, variances.ctypes.data))
        return means, variances

    # --- all creatures -------------------------------------------------------------------------
    def iterate(self):
        check(lib().cn_iterate(self._h))

    def iterate_and_render(self):
        """One step; its diffusion pass also packs channels 0..2 into each creature's frame."""
        check(lib().cn_iterate_and_render(self._h))

    def run_steps(self, n):
        check(lib().cn_run_steps(self._h, int(n)))

    def sync(self):
        check(lib().cn_sync(self._h))

    def iterate_on_stream(self, stream):
        """One step enqueued on a caller's HIP stream (torch.cuda.Stream, a raw handle, or None
        for the null stream), ordered after the engine's queued work (iterate_with_encoder,
        mod.rs:435-466)."""
        handle = getattr(stream, "cuda_stream", stream)
        check(lib().cn_iterate_on_stream(self._h, ctypes.c_void_p(handle or 0)))

    def stream_handle(self):
        """The engine's own hipStream_t as an integer (wrap with torch.cuda.ExternalStream)."""
        h = ctypes.c_void_p()
        check(lib().cn_engine_stream(self._h, ctypes.byref(h)))
        return h.value or 0

    def creature_state_ptr(self, i):
        """(device address, element count) of creature i's current state (channel_buffer,
        mod.rs:202-204); valid until the next step or rollout."""
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        check(lib().cn_creature_state_ptr(self._h, int(i), ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def rollout_grad(self, steps, ckpt_every):
        """Forward `steps` steps, MSE vs each creature's target, checkpointed adjoint (async)."""
        check(lib().cn_rollout_grad(self._h, int(steps), int(ckpt_every)))

    def rollout_forward(self, steps, ckpt_every):
        check(lib().cn_rollout_forward(self._h, int(steps), int(ckpt_every)))

    def rollout_backward(self, grad_finals):
        """dL/dx_T per creature as contiguous f32 device tensors."""
        if len(grad_finals) != len(self._shapes):
            raise ValueError(f"expected {len(self._shapes)} gradients, got {len(grad_finals)}")
        arr = (ctypes.c_void_p * len(grad_finals))(
            *[_dev_f32(g, self._elems(i), f"grad_finals[{i}]") for i, g in enumerate(grad_finals)])
        check(lib().cn_rollout_backward(self._h, arr, len(grad_finals)))

    # --- creature search on the device ----------------------------------------------------
    def _nparams(self):
        return sum(3 * K for _, _, K in self._shapes)

    def batch_growth(self):
        """All creatures' (mu, sigma, w): a flat float32 array, creature i's [mu K | sigma K | w K]
        after creature i-1's."""
        out = np.empty(self._nparams(), np.float32)
        check(lib().cn_batch_get_growth(self._h, out.ctypes.data, out.size))
        return out

    def batch_set_growth(self, params):
        a = _f32(params).reshape(-1)
        check(lib().cn_batch_set_growth(self._h, a.ctypes.data, a.size))

    def batch_param_grads(self):
        """dL/d(mu, sigma, w) of the last backward for all creatures, laid out as batch_growth."""
        out = np.empty(self._nparams(), np.float32)
        check(lib().cn_batch_get_param_grads(self._h, out.ctypes.data, out.size))
        return out

    def adam_step(self, lr=5e-3, beta1=0.9, beta2=0.999, eps=1e-8, mu_min=0.0, mu_max=2.0, sigma_min=1e-3,
                  weight_min=0.0):
        """Device-side Adam on every creature's (mu, sigma, w) from the last backward, then the
        sanitize box of the reference's LeniaParams (asynchronous)."""
        p = AdamParams(lr, beta1, beta2, eps, mu_min, mu_max, sigma_min, weight_min)
        check(lib().cn_adam_step(self._h, ctypes.byref(p)))

    def adam_reset(self):
        check(lib().cn_adam_reset(self._h))

    def autotune(self, reps=5):
        """Pick the fastest FFT / stencil geometry per grid size on the device."""
        check(lib().cn_autotune(self._h, int(reps)))
        return self.plans()

    def plans(self):
        out = []
        for i in range(lib().cn_engine_size_classes(self._h)):
            p = Plan()
            check(lib().cn_engine_get_plan(self._h, i, ctypes.byref(p)))
            out.append({f: getattr(p, f) for f, _ in Plan._fields_})
        return out

    def set_plan(self, n, fwd_e=0, inv_e=0, col_e=0, col_tc=0, diff_th=0, col_split=0, col_full=0):
        p = Plan(int(n), int(fwd_e), int(inv_e), int(col_e), int(col_tc), int(diff_th), int(col_split), int(col_full))
        check(lib().cn_engine_set_plan(self._h, ctypes.byref(p)))

    def info(self):
        """Engine summary; per-creature fields (n, channels, kernels, compact kernels, segments)
        describe creature 0."""
        i = EngineInfo()
        check(lib().cn_engine_get_info(self._h, ctypes.byref(i)))
        return {"n": i.n, "num_channels": i.num_channels, "num_kernels": i.num_kernels,
                "steps_done": i.steps_done, "device_bytes": i.device_bytes, "compact_kernels": i.compact_kernels,
                "segments": i.segments, "segment_radius": i.segment_radius, "segment_length": i.segment_length}

    def profile_kernels(self, steps):
        """Average device ms per launch of each pass of the first size class over `steps` of
        4 * steps direct steps queued back to back (events around one pass per step): [row
        forward, column pass, inverse row + growth, diffusion].  The state advances 4 * steps."""
        out = np.empty(4, dtype=np.float64)
        check(lib().cn_profile_kernels(self._h, int(steps), out.ctypes.data, out.size))
        return out


class GpuMaceLenia(_Engine):
    """GpuMaceLenia::new (mod.rs:79-198): square grid shape[0] x shape[0], C channels, K kernels,
    c0 = input channel of each kernel, c1 = output channel, growth mu/sigma/weights, temp."""

    def __init__(self, shape, num_channels, num_kernels, c0, c1, mu, sigma, weights, temp=1.0, device=0):
        if isinstance(shape, int):
            shape = (shape, shape)
        if len(shape) != 2:
            raise ValueError("DiffusionLenia requires 2D grids")  # mod.rs:94
        n = int(shape[0])
        self._keep = []
        self._init([_desc(n, num_channels, num_kernels, c0, c1, mu, sigma, weights, temp, device, self._keep)],
                   device)
        self.n, self.num_channels, self.num_kernels = n, int(num_channels), int(num_kernels)
        self.total_elements = n * n

    # --- uploads (mod.rs:207-226) ---------------------------------------------------------
    def set_kernel(self, kernel_fft, kernel_idx):
        """Full N*N complex64 spectrum of kernel `kernel_idx` (complex array or interleaved f32)."""
        self.creature_set_kernel(0, kernel_idx, kernel_fft)

    def upload_channel(self, data, channel):
        """f64 values cast to f32 on upload (mod.rs:220); f32 input is uploaded as is."""
        a = np.asarray(data)
        if a.dtype == np.float32:
            self.creature_upload_channel(0, channel, a)
        else:
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            check(lib().cn_upload_channel_f64(self._h, int(channel), a.ctypes.data, a.size))

    def set_growth(self, mu, sigma, weights):
        self.creature_set_growth(0, mu, sigma, weights)

    # --- stepping (mod.rs:229-233, 348-386, 469-540) --------------------------------------
    def timed_iterate(self):
        """[conv, growth, diffusion, copy] device milliseconds of one step."""
        ms = (ctypes.c_double * 4)()
        check(lib().cn_timed_iterate(self._h, ms))
        return [ms[i] for i in range(4)]

    # --- downloads (mod.rs:236-299) -------------------------------------------------------
    def download_all_channels(self):
        return self.creature_download(0)

    def download_conv_buffer(self):
        """Real part of the unnormalised convolution buffer, K*N*N f32 (mod.rs:267-299)."""
        out = np.empty(self.num_kernels * self.total_elements, dtype=np.float32)
        check(lib().cn_download_conv_real(self._h, out.ctypes.data, out.size))
        return out

    def download_render(self):
        """RGB24 bytes of the last rendered frame, N*N*3 uint8 (mod.rs:302-335)."""
        return self.creature_download_render(0).reshape(-1)

    def render_features(self):
        """Audio features of the last rendered frame, computed on the device."""
        return self.creature_render_features(0)

    def mass_fp64(self):
        """Per-channel FP64 compensated mass of the current state."""
        return self.creature_mass(0)



class CreatureBatch(_Engine):
    """Many creatures (mixed N, C, K) advanced together: one launch per kernel per grid size,
    all captured in one HIP graph per step / per rollout."""

    def __init__(self, creatures, device=0):
        self._keep = []
        descs = [_desc(c["n"], c["num_channels"], c["num_kernels"], c["src"], c["dst"], c["mu"], c["sigma"],
                       c["weights"], c.get("temp", 1.0), device, self._keep) for c in creatures]
        self._init(descs, device)
        self.count = len(creatures)

    @classmethod
    def generated(cls, specs, device=0):
        """A batch whose creatures (creatures.creature_spec dicts) are generated on the device:
        no host FFTs and no spectrum uploads."""
        placeholder = [{**sp, "mu": np.full(sp["num_kernels"], 0.1, np.float32),
                        "sigma": np.full(sp["num_kernels"], 0.05, np.float32),
                        "weights": np.full(sp["num_kernels"], 1.0, np.float32)} for sp in specs]
        b = cls(placeholder, device)
        for i, sp in enumerate(specs):
            b.creature_generate(i, sp["radius"], sp.get("patch", 128), sp.get("seed", 0), sp.get("creature", 0))
        return b

    def load(self, i, cr):
        """Upload a generated creature's kernel spectra and seed (creatures.generate_creature)."""
        for k in range(cr["num_kernels"]):
            self.creature_set_kernel(i, k, cr["kernels_fft"][k])
        for c in range(cr["num_channels"]):
            self.creature_upload_channel(i, c, cr["state"][c])
