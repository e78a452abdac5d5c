"""ORACLE - test infrastructure only.  Imported by tests/, __graft_entry__.smoke() and the CPU
baseline leg of bench.py, as the checker; the product never imports it.

Two independent CPU restatements of one DiffusionLenia step:
  * liboracle.so (cpu_ref.cpp): the reference's CPU implementation ml-rs/example.rs:42-303 in
    f32 (the "ml-rs-example" baseline) and FP64;
  * step_numpy_f64: the PyTorch reference's semantics (train/diff_lenia_org.py:27-53,
    train/lenia_org.py:232-275) with numpy.fft in FP64, used where the C++ FP64 oracle is too
    slow (N >= 1024);
plus brute-force pins that share no code with either: a matrix DFT and a direct circular
convolution in FP64; and the render epilogue's colour pack (render_rgb).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_SIGS = {
    "oracle_run_f32": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, _P, _P, ctypes.c_float,
                       ctypes.c_int],
    "oracle_run_f64": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, _P, _P, ctypes.c_float,
                       ctypes.c_int],
    "oracle_conv_f64": [ctypes.c_int, ctypes.c_int, _P, _P, _P, _P],
    "oracle_conv_f32": [ctypes.c_int, ctypes.c_int, _P, _P, _P, _P],
    "oracle_growth_f64": [ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_float, ctypes.c_int, _P],
    "oracle_diffusion_f64": [ctypes.c_int, _P, _P, _P],
    "oracle_fft2_f64": [ctypes.c_int, _P],
    "oracle_fft2_f32": [ctypes.c_int, _P],
    "oracle_run_f32_threads": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, _P, _P, ctypes.c_float,
                               ctypes.c_int, ctypes.c_int],
    "oracle_example_creature_f32": [ctypes.c_int, _P, _P, _P, _P, _P, _P, _P],
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            getattr(h, name).argtypes = args
            getattr(h, name).restype = None
        _lib = h
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(np.asarray(a, dtype=dt))


def _kf(kfft, K, n):
    a = np.ascontiguousarray(np.asarray(kfft).reshape(K, n * n).astype(np.complex64))
    return a.view(np.float32)


def run_f32(state, kfft, src, dst, mu, sigma, w, temp=1.0, steps=1, threads=1):
    """`steps` steps of ml-rs-example's f32 arithmetic; state [C,n,n] -> new f32 array.  threads > 1
    shares the kernels' convolutions out over threads (the same results as one thread)."""
    st = _c(state, np.float32).copy()
    C, n, _ = st.shape
    K = len(src)
    kf, s_, d_ = _kf(kfft, K, n), _c(src, np.uint32), _c(dst, np.uint32)
    m, sg, ww = _c(mu, np.float32), _c(sigma, np.float32), _c(w, np.float32)
    if threads > 1:
        lib().oracle_run_f32_threads(n, C, K, st.ctypes.data, kf.ctypes.data, s_.ctypes.data, d_.ctypes.data,
                                     m.ctypes.data, sg.ctypes.data, ww.ctypes.data, float(temp), int(steps),
                                     int(threads))
    else:
        lib().oracle_run_f32(n, C, K, st.ctypes.data, kf.ctypes.data, s_.ctypes.data, d_.ctypes.data, m.ctypes.data,
                             sg.ctypes.data, ww.ctypes.data, float(temp), int(steps))
    return st


def example_creature_f32(n=1024):
    """The reference CPU example's creature in its own f32 arithmetic (example.rs:310-491): dict with
    kernels_fft (9, n*n) complex64, state (3, n, n) f32, mu/sigma/weights (9,) f32, src/dst (9,)."""
    K, C, P = 9, 3, n * n
    kf = np.empty(2 * K * P, np.float32)
    st = np.empty(C * P, np.float32)
    mu, sg, w = (np.empty(K, np.float32) for _ in range(3))
    c0, c1 = np.empty(K, np.uint32), np.empty(K, np.uint32)
    lib().oracle_example_creature_f32(n, kf.ctypes.data, st.ctypes.data, mu.ctypes.data, sg.ctypes.data,
                                      w.ctypes.data, c0.ctypes.data, c1.ctypes.data)
    return {"n": n, "num_channels": C, "num_kernels": K, "kernels_fft": kf.view(np.complex64).reshape(K, P),
            "state": st.reshape(C, n, n), "mu": mu, "sigma": sg, "weights": w, "src": c0, "dst": c1}


def run_f64(state, kfft, src, dst, mu, sigma, w, temp=1.0, steps=1):
    """Same algorithm in FP64 (the FP64 oracle); state [C,n,n] -> new f64 array."""
    st = _c(state, np.float64).copy()
    C, n, _ = st.shape
    K = len(src)
    kf, s_, d_ = _kf(kfft, K, n), _c(src, np.uint32), _c(dst, np.uint32)
    m, sg, ww = _c(mu, np.float32), _c(sigma, np.float32), _c(w, np.float32)
    lib().oracle_run_f64(n, C, K, st.ctypes.data, kf.ctypes.data, s_.ctypes.data, d_.ctypes.data, m.ctypes.data,
                         sg.ctypes.data, ww.ctypes.data, float(temp), int(steps))
    return st


def conv_f64(state, kfft, src):
    st = _c(state, np.float64)
    C, n, _ = st.shape
    K = len(src)
    out = np.empty((K, n, n), np.float64)
    kf, s_ = _kf(kfft, K, n), _c(src, np.uint32)
    lib().oracle_conv_f64(n, K, st.ctypes.data, kf.ctypes.data, s_.ctypes.data, out.ctypes.data)
    return out


def conv_f32(state, kfft, src):
    st = _c(state, np.float32)
    C, n, _ = st.shape
    K = len(src)
    out = np.empty((K, n, n), np.float32)
    kf, s_ = _kf(kfft, K, n), _c(src, np.uint32)
    lib().oracle_conv_f32(n, K, st.ctypes.data, kf.ctypes.data, s_.ctypes.data, out.ctypes.data)
    return out


def growth_f64(u, dst, mu, sigma, w, temp, out_ch):
    uu = _c(u, np.float64)
    K, n, _ = uu.shape
    out = np.empty((n, n), np.float64)
    d_, m, sg, ww = _c(dst, np.uint32), _c(mu, np.float32), _c(sigma, np.float32), _c(w, np.float32)
    lib().oracle_growth_f64(n, K, uu.ctypes.data, d_.ctypes.data, m.ctypes.data, sg.ctypes.data, ww.ctypes.data,
                            float(temp), int(out_ch), out.ctypes.data)
    return out


def diffusion_f64(affinity, channel):
    a, ch = _c(affinity, np.float64), _c(channel, np.float64)
    n = a.shape[0]
    out = np.empty((n, n), np.float64)
    lib().oracle_diffusion_f64(n, a.ctypes.data, ch.ctypes.data, out.ctypes.data)
    return out


def fft2_f64(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    lib().oracle_fft2_f64(a.shape[0], a.ctypes.data)
    return a


def fft2_f32(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex64)).copy()
    lib().oracle_fft2_f32(a.shape[0], a.ctypes.data)
    return a


# ----------------------------------------------------------------------------------------
# numpy FP64 restatement of the PyTorch reference step (diff_lenia_org.py:27-53)
# ----------------------------------------------------------------------------------------
def box3(x):
    """3x3 periodic box sum over the last two axes (F.pad circular + unfold + sum)."""
    s = np.zeros_like(x)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            s += np.roll(np.roll(x, dy, axis=-2), dx, axis=-1)
    return s


def conv_numpy_f64(state, kfft, src):
    """u_k = Re(ifft2(fft2(state[src_k]) * K_k)) in FP64 (lenia_org.py:266-275)."""
    st = np.asarray(state, dtype=np.float64)
    C, n, _ = st.shape
    K = len(src)
    S = np.fft.fft2(st)
    kf = np.asarray(kfft).reshape(K, n, n).astype(np.complex128)
    return np.stack([np.real(np.fft.ifft2(S[int(src[k])] * kf[k])) for k in range(K)])


def affinity_numpy_f64(u, dst, mu, sigma, w, temp, C):
    """A_c = exp(temp * sum_{k: dst_k = c} w_k * (2 exp(-((u_k - mu_k)/sigma_k)^2 / 2) - 1))."""
    K, n, _ = u.shape
    aff = np.zeros((C, n, n))
    for k in range(K):
        m, s = float(np.float32(mu[k])), float(np.float32(sigma[k]))
        g = 2.0 * np.exp(-((u[k] - m) ** 2 / s ** 2) / 2.0) - 1.0
        aff[int(dst[k])] += float(np.float32(w[k])) * g
    return np.exp(float(temp) * aff)


def diffusion_numpy_f64(A, state):
    Z = box3(A)
    return A * box3(state / Z)


def step_numpy_f64(state, kfft, src, dst, mu, sigma, w, temp=1.0):
    st = np.asarray(state, dtype=np.float64)
    u = conv_numpy_f64(st, kfft, src)
    A = affinity_numpy_f64(u, dst, mu, sigma, w, temp, st.shape[0])
    return diffusion_numpy_f64(A, st)


# ----------------------------------------------------------------------------------------
# brute-force pins (no FFT anywhere)
# ----------------------------------------------------------------------------------------
def dft_matrix(n, inverse=False):
    j = np.arange(n)
    sgn = 1.0 if inverse else -1.0
    return np.exp(sgn * 2j * np.pi * np.outer(j, j) / n)


def dft2_direct(x):
    F = dft_matrix(x.shape[0])
    return F @ np.asarray(x, dtype=np.complex128) @ F


def idft2_direct(x):
    n = x.shape[0]
    F = dft_matrix(n, inverse=True)
    return F @ np.asarray(x, dtype=np.complex128) @ F / (n * n)


def conv_direct(channel, kernel):
    """Circular convolution u[y,x] = sum_{a,b} channel[y-a, x-b] kernel[a,b] (FP64, O(n^4))."""
    ch = np.asarray(channel, dtype=np.float64)
    k = np.asarray(kernel, dtype=np.float64)
    out = np.zeros_like(ch)
    n = ch.shape[0]
    for a in range(n):
        for b in range(n):
            if k[a, b] != 0.0:
                out += k[a, b] * np.roll(np.roll(ch, a, axis=0), b, axis=1)
    return out


def render_rgb(state):
    """RGB24 frame of channels 0..2 (render.rs:145-153 dispatch, compute_1024.wgsl:429-443 and the
    unpacking of mod.rs:325-330): byte = u32(sqrt(clamp(1.5 c, 0, 1)) * 255) in f32, IEEE sqrt,
    truncation toward zero; channels beyond C read as 0.  Returns (N, N, 3) uint8."""
    st = np.asarray(state, dtype=np.float32)
    C, n = st.shape[0], st.shape[1]
    out = np.zeros((n, n, 3), dtype=np.uint8)
    for c in range(min(C, 3)):
        x = np.clip(st[c] * np.float32(1.5), np.float32(0.0), np.float32(1.0))
        out[..., c] = (np.sqrt(x) * np.float32(255.0)).astype(np.float32).astype(np.uint32).astype(np.uint8)
    return out


def rel_linf(x, ref):
    """max|x - ref| / max|ref| over all entries."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    x = x.astype(np.complex128) if np.iscomplexobj(x) or np.iscomplexobj(ref) else x.astype(np.float64)
    ref = ref.astype(x.dtype)
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300))
