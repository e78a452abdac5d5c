"""Synthetic benchmark creatures (C ABI cn_generate_creature): Lenia ring kernels of radius R with
1-3 rings, growth mu ~ U[0.05,0.15], sigma ~ U[0.03,0.06], Dirichlet(1) weights per destination,
U[0,1) seed on the central patch.  Deterministic in (seed, creature)."""
import ctypes

import numpy as np

from ._lib import CreatureParams, check, lib


def channel_maps(num_channels, num_kernels=None):
    """All src x dst pairs in the reference's ordering: c0[k] = k % C (input), c1[k] = k // C
    (output) (ml-rs/example.rs:473-477)."""
    K = num_channels * num_channels if num_kernels is None else num_kernels
    k = np.arange(K, dtype=np.uint32)
    return k % num_channels, (k // num_channels) % num_channels


def creature_params(n, num_channels, num_kernels, src, dst, radius, patch=128, seed=0, creature=0):
    """cn_creature_params for the C ABI; returns (params, keep-alive arrays)."""
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    rad = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, dtype=np.uint32), (num_kernels,)))
    p = CreatureParams(n, num_channels, num_kernels,
                       src.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                       dst.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                       rad.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                       int(patch), int(seed), int(creature))
    return p, (src, dst, rad)


def generate_creature(n, num_channels, num_kernels, src, dst, radius, patch=128, seed=0, creature=0,
                      spectra=True, spatial=False, state=True):
    p, (src, dst, rad) = creature_params(n, num_channels, num_kernels, src, dst, radius, patch, seed, creature)
    P = n * n
    kf = np.empty(2 * num_kernels * P, np.float32) if spectra else None
    sp = np.empty(num_kernels * P, np.float64) if spatial else None
    st = np.empty(num_channels * P, np.float32) if state else None
    mu = np.empty(num_kernels, np.float32)
    sg = np.empty(num_kernels, np.float32)
    w = np.empty(num_kernels, np.float32)

    def ptr(a):
        return a.ctypes.data if a is not None else None

    check(lib().cn_generate_creature(ctypes.byref(p), ptr(kf), ptr(sp), ptr(st), mu.ctypes.data,
                                     sg.ctypes.data, w.ctypes.data))
    out = {"n": n, "num_channels": num_channels, "num_kernels": num_kernels, "src": src, "dst": dst,
           "radius": rad, "mu": mu, "sigma": sg, "weights": w}
    if spectra:
        out["kernels_fft"] = kf.view(np.complex64).reshape(num_kernels, P)
    if spatial:
        out["spatial"] = sp.reshape(num_kernels, n, n)
    if state:
        out["state"] = st.reshape(num_channels, n, n)
    return out


def baseline_creature(n, num_channels=3, radius=13, seed=0, creature=0, patch=128, **kw):
    """K = C^2 creature with the reference's c0/c1 maps (e.g. 1024^2 C3K9 R=13; 2048^2 R=26)."""
    src, dst = channel_maps(num_channels)
    return generate_creature(n, num_channels, len(src), src, dst, radius, patch=patch, seed=seed,
                             creature=creature, **kw)


def make_engine(cr, temp=1.0, device=0):
    """GpuMaceLenia engine loaded with a generated creature's kernels and seed."""
    from .engine import GpuMaceLenia
    e = GpuMaceLenia(cr["n"], cr["num_channels"], cr["num_kernels"], cr["src"], cr["dst"], cr["mu"],
                     cr["sigma"], cr["weights"], temp=temp, device=device)
    for k in range(cr["num_kernels"]):
        e.set_kernel(cr["kernels_fft"][k], k)
    for c in range(cr["num_channels"]):
        e.upload_channel(cr["state"][c], c)
    return e


def creature_spec(n, num_channels, radius, patch=128, seed=0, creature=0, src=None, dst=None):
    """Shape and generator keys of a creature (no data): K = C^2 with the reference maps unless
    src/dst are given.  CreatureBatch.generated builds and fills creatures from these on the
    device."""
    if src is None:
        src, dst = channel_maps(num_channels)
    K = len(src)
    return {"n": int(n), "num_channels": int(num_channels), "num_kernels": K,
            "src": np.asarray(src, np.uint32), "dst": np.asarray(dst, np.uint32),
            "radius": np.broadcast_to(np.asarray(radius, np.uint32), (K,)).copy(), "patch": int(patch),
            "seed": int(seed), "creature": int(creature)}


def uneven_batch_specs(seed=0, sizes=None):
    """The uneven creature batch: 85 creatures at 256^2, 85 at 512^2 and 86 at 1024^2 (in that
    order), channels C ~ U{2,3,4} with K = C^2 (all src x dst pairs), radius R ~ U{8..20}, drawn in
    creature order from numpy's default_rng(seed); creature i uses generator keys (seed, i)."""
    rng = np.random.default_rng(seed)
    sizes = [256] * 85 + [512] * 85 + [1024] * 86 if sizes is None else list(sizes)
    specs = []
    for i, n in enumerate(sizes):
        C = int(rng.integers(2, 5))
        specs.append(creature_spec(n, C, int(rng.integers(8, 21)), seed=seed, creature=i))
    return specs


def search_batch_creatures(count=32, n=512, num_channels=5, seed=0):
    """The batched-search creatures: `count` creatures of n^2, C channels, K = C^2 kernels, radius
    R ~ U{10..18} (numpy default_rng(seed)), seed patch 128^2, and their targets: creature
    1000 + i of the same distribution (to be run forward T steps by the caller)."""
    rng = np.random.default_rng(seed)
    radii = rng.integers(10, 19, size=count)
    crs = [baseline_creature(n, num_channels=num_channels, radius=int(radii[i]), seed=seed, creature=i)
           for i in range(count)]
    tgs = [baseline_creature(n, num_channels=num_channels, radius=int(radii[i]), seed=seed, creature=1000 + i)
           for i in range(count)]
    return crs, tgs


def example_creature(n=1024, spectra=True, spatial=False, state=True):
    """The creature of the reference's CPU example, ml-rs-example (ml-rs/example.rs:310-491):
    C = 3, K = 9, c0 = k % 3 (input), c1 = k // 3 (output), sigmoid x Gaussian ring kernels,
    Gaussian-blob seed, the example's growth parameters; at grid size n (the example uses 1024)."""
    K, C, P = 9, 3, n * n
    kf = np.empty(2 * K * P, np.float32) if spectra else None
    sp = np.empty(K * P, np.float64) if spatial else None
    st = np.empty(C * P, np.float32) if state else None
    mu, sg, w = (np.empty(K, np.float32) for _ in range(3))
    src, dst = np.empty(K, np.uint32), np.empty(K, np.uint32)

    def ptr(a):
        return a.ctypes.data if a is not None else None

    check(lib().cn_generate_example_creature(n, ptr(kf), ptr(sp), ptr(st), mu.ctypes.data, sg.ctypes.data,
                                             w.ctypes.data, src.ctypes.data, dst.ctypes.data))
    out = {"n": n, "num_channels": C, "num_kernels": K, "src": src, "dst": dst, "mu": mu, "sigma": sg, "weights": w}
    if spectra:
        out["kernels_fft"] = kf.view(np.complex64).reshape(K, P)
    if spatial:
        out["spatial"] = sp.reshape(K, n, n)
    if state:
        out["state"] = st.reshape(C, n, n)
    return out
