"""Python mirror of the reference orchestrator `GpuMaceLenia` (ml-rs/orchestrator/mod.rs:45-541)
over the C ABI.  Same method set and argument meaning; errors raise CellaNovaError instead of
panicking (mod.rs:94,208,218)."""
import ctypes

import numpy as np

from ._lib import EngineDesc, EngineInfo, check, lib


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class GpuMaceLenia:
    """GpuMaceLenia::new (mod.rs:79-198): square grid shape[0] x shape[0], C channels, K kernels,
    c0 = input channel of each kernel, c1 = output channel, growth mu/sigma/weights, temp."""

    def __init__(self, shape, num_channels, num_kernels, c0, c1, mu, sigma, weights, temp=1.0, device=0):
        if isinstance(shape, int):
            shape = (shape, shape)
        if len(shape) != 2:
            raise ValueError("DiffusionLenia requires 2D grids")  # mod.rs:94
        n = int(shape[0])
        self._keep = [_u32(c0), _u32(c1), _f32(mu), _f32(sigma), _f32(weights)]
        for a in self._keep:
            if a.shape != (num_kernels,):
                raise ValueError(f"per-kernel arrays must have {num_kernels} entries")
        src, dst, m, s, w = self._keep
        desc = EngineDesc(
            n, num_channels, num_kernels,
            src.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            dst.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            s.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            float(temp), int(device))
        h = ctypes.c_void_p()
        check(lib().cn_engine_create(ctypes.byref(desc), ctypes.byref(h)))
        self._h = h
        self.n, self.num_channels, self.num_kernels = n, int(num_channels), int(num_kernels)
        self.total_elements = n * n

    # --- lifecycle -------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
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

    # --- uploads (mod.rs:207-226) ---------------------------------------------------------
    def set_kernel(self, kernel_fft, kernel_idx):
        """Full N*N complex64 spectrum of kernel `kernel_idx` (complex array or interleaved f32)."""
        a = np.asarray(kernel_fft)
        if np.iscomplexobj(a):
            a = np.ascontiguousarray(a.astype(np.complex64)).view(np.float32)
        else:
            a = _f32(a)
        a = a.reshape(-1)
        check(lib().cn_set_kernel(self._h, int(kernel_idx), a.ctypes.data, a.size // 2))

    def upload_channel(self, data, channel):
        """f64 values cast to f32 on upload (mod.rs:220); f32 input is uploaded as is."""
        a = np.asarray(data)
        if a.dtype == np.float32:
            a = np.ascontiguousarray(a).reshape(-1)
            check(lib().cn_upload_channel_f32(self._h, int(channel), a.ctypes.data, a.size))
        else:
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            check(lib().cn_upload_channel_f64(self._h, int(channel), a.ctypes.data, a.size))

    # --- stepping (mod.rs:229-233, 348-386, 469-540) --------------------------------------
    def iterate(self):
        check(lib().cn_iterate(self._h))

    def run_steps(self, n):
        check(lib().cn_run_steps(self._h, int(n)))

    def sync(self):
        check(lib().cn_sync(self._h))

    def timed_iterate(self):
        """[conv, growth, diffusion, copy] device milliseconds of one step."""
        ms = (ctypes.c_double * 4)()
        check(lib().cn_timed_iterate(self._h, ms))
        return [ms[i] for i in range(4)]

    def profile_kernels(self, steps):
        """Average device ms per launch over `steps` directly launched steps:
        [row forward, fused column, inverse row + growth, diffusion]."""
        out = np.empty(4, dtype=np.float64)
        check(lib().cn_profile_kernels(self._h, int(steps), out.ctypes.data, out.size))
        return out

    # --- downloads (mod.rs:236-299) -------------------------------------------------------
    def download_all_channels(self):
        out = np.empty(self.num_channels * self.total_elements, dtype=np.float32)
        check(lib().cn_download_channels(self._h, out.ctypes.data, out.size))
        return out

    def download_conv_buffer(self):
        """Real part of the unnormalised convolution buffer, K*N*N f32 (mod.rs:267-299)."""
        out = np.empty(self.num_kernels * self.total_elements, dtype=np.float32)
        check(lib().cn_download_conv_real(self._h, out.ctypes.data, out.size))
        return out

    def mass_fp64(self):
        """Per-channel FP64 compensated mass of the current state."""
        out = np.empty(self.num_channels, dtype=np.float64)
        check(lib().cn_mass_fp64(self._h, out.ctypes.data, out.size))
        return out

    def info(self):
        i = EngineInfo()
        check(lib().cn_engine_get_info(self._h, ctypes.byref(i)))
        return {"n": i.n, "num_channels": i.num_channels, "num_kernels": i.num_kernels,
                "steps_done": i.steps_done, "device_bytes": i.device_bytes}
