"""Loader mirror of ml-rs/config.rs: SimConfig, load_config (with the c0/c1 swap, config.rs:40)
and load_kernels (with spectral zero-padding, config.rs:56-125), and the writers of the same two
files (train/generate_kernel_json.py:171-213), implemented in the C ABI."""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import Config, check, lib


@dataclass
class SimConfig:
    """config.rs:14-25, after the swap: c0 is each kernel's input channel, c1 its output."""
    seed_size: int
    num_channels: int
    num_kernels: int
    seed_channels: list
    c0: np.ndarray
    c1: np.ndarray
    growth_mu: np.ndarray
    growth_sigma: np.ndarray
    growth_weights: np.ndarray


def load_config(path):
    ptr = ctypes.POINTER(Config)()
    check(lib().cn_load_config(str(path).encode(), ctypes.byref(ptr)))
    try:
        c = ptr.contents
        C, K, P = c.num_channels, c.num_kernels, c.seed_len

        def arr(p, n, dt):
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True) if n else np.zeros(0, dt)

        seeds = arr(c.seed_channels, C * P, np.float64).reshape(C, P) if C * P else np.zeros((C, 0))
        return SimConfig(
            seed_size=c.seed_size, num_channels=C, num_kernels=K,
            seed_channels=[seeds[i].copy() for i in range(C)],
            c0=arr(c.c0, K, np.uint32), c1=arr(c.c1, K, np.uint32),
            growth_mu=arr(c.growth_mu, K, np.float32), growth_sigma=arr(c.growth_sigma, K, np.float32),
            growth_weights=arr(c.growth_weights, K, np.float32))
    finally:
        lib().cn_free_config(ptr)


def load_kernels(path, num_kernels, target_size):
    """K spectra as complex64 [K, target_size*target_size]."""
    out = np.empty(2 * num_kernels * target_size * target_size, dtype=np.float32)
    check(lib().cn_load_kernels(str(path).encode(), int(num_kernels), int(target_size), out.ctypes.data, out.size))
    return out.view(np.complex64).reshape(num_kernels, target_size * target_size)


def write_kernels(path, spectra):
    """K spectra (K, n*n) complex64 -> .bin of little-endian (re, im) f32 pairs (config.rs:49-55)."""
    a = np.ascontiguousarray(np.asarray(spectra, dtype=np.complex64))
    K = a.shape[0]
    n = int(round(np.sqrt(a[0].size)))
    if n * n != a[0].size:
        raise ValueError("each spectrum must hold n*n values")
    f = a.reshape(-1).view(np.float32)
    check(lib().cn_write_kernels(str(path).encode(), f.ctypes.data, K, n))


def write_config(path, cfg):
    """SimConfig (loader convention: c0 = input channel) -> creature JSON that load_config reads
    back to the same values."""
    C, K = int(cfg.num_channels), int(cfg.num_kernels)
    seeds = np.ascontiguousarray(np.asarray(cfg.seed_channels, dtype=np.float64).reshape(C, -1))
    c0, c1 = (np.ascontiguousarray(np.asarray(a, np.uint32)) for a in (cfg.c0, cfg.c1))
    mu, sg, w = (np.ascontiguousarray(np.asarray(a, np.float32))
                 for a in (cfg.growth_mu, cfg.growth_sigma, cfg.growth_weights))
    P = ctypes.POINTER
    c = Config(int(cfg.seed_size), C, K, seeds.ctypes.data_as(P(ctypes.c_double)), seeds.shape[1],
               c0.ctypes.data_as(P(ctypes.c_uint32)), c1.ctypes.data_as(P(ctypes.c_uint32)),
               mu.ctypes.data_as(P(ctypes.c_float)), sg.ctypes.data_as(P(ctypes.c_float)),
               w.ctypes.data_as(P(ctypes.c_float)))
    check(lib().cn_write_config(str(path).encode(), ctypes.byref(c)))
