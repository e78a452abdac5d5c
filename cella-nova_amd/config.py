"""Loader mirror of ml-rs/config.rs: SimConfig, load_config (with the c0/c1 swap, config.rs:40)
and load_kernels (with spectral zero-padding, config.rs:56-125), implemented in the C ABI."""
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
