"""Writes the committed golden fixtures tests/golden/step_*.npz and example_n16.npz.

Inputs are small synthetic creatures (seeded generator of the package).  Expected outputs are
computed by brute force in FP64 with no FFT anywhere: the effective spatial kernel is the real
part of a matrix inverse DFT of the complex64 spectrum, the convolution is a direct circular
sum, then growth, affinity, periodic 3x3 Z and the mass-conserving gather follow the reference
step (ml-rs/example.rs:164-303, train/diff_lenia_org.py:27-53).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def brute_step(state, kfft, src, dst, mu, sigma, w, temp):
    orc = ge.load_oracle()
    C, n, _ = state.shape
    K = len(src)
    kern = [np.real(orc.idft2_direct(np.asarray(kfft[k], np.complex128).reshape(n, n))) for k in range(K)]
    u = np.stack([orc.conv_direct(state[int(src[k])], kern[k]) for k in range(K)])
    aff = np.zeros((C, n, n))
    for k in range(K):
        m, s, wk = float(mu[k]), float(sigma[k]), float(w[k])
        aff[int(dst[k])] += wk * (2.0 * np.exp(-0.5 * (u[k] - m) ** 2 / s ** 2) - 1.0)
    A = np.exp(float(np.float32(temp)) * aff)  # the engine takes temp as f32
    Z = orc.box3(A)
    return A * orc.box3(state / Z), u


def example_creature_brute(n):
    """The reference CPU example's creature (ml-rs/example.rs:310-491) restated in FP64 numpy
    from its formulas, with a brute-force matrix DFT: spatial kernels (centre (n/2, n/2),
    sigmoid x Gaussian ring, sum 1), the ifftshift as an explicit index map (output (y, x)
    takes the centred value at ((y + n/2) mod n, (x + n/2) mod n)), seed blobs, growth
    parameters in float32 arithmetic."""
    orc = ge.load_oracle()
    radii = [0.25, 0.3167, 0.3833, 0.45, 0.5167, 0.5833, 0.65, 0.7167, 0.85]
    widths = [0.03, 0.0411, 0.0522, 0.0633, 0.0744, 0.0856, 0.0967, 0.1078, 0.13]
    h = n // 2
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    spatial, spectra = [], []
    for r, w in zip(radii, widths):
        d = np.sqrt((ii - h) ** 2.0 + (jj - h) ** 2.0) / (10.0 * r)
        cen = (0.5 * np.tanh(-(d - 1.0) * 10.0 * 0.5) + 0.5) * np.exp(-(d - 0.5) ** 2 / (2.0 * w * w))
        cen = cen / cen.sum()
        shifted = cen[(ii + h) % n, (jj + h) % n]
        spatial.append(shifted)
        spectra.append(orc.dft2_direct(shifted.astype(np.complex128)))
    x, y = (jj - h) / h, (ii - h) / h
    seed = np.stack([np.exp(-((x - ox) ** 2 + (y - oy) ** 2) / (2.0 * s * s))
                     for s, ox, oy in ((0.25, 0.0, 0.0), (0.22, 0.04, 0.0), (0.28, 0.0, 0.04))])
    f = np.float32
    k = np.arange(9)
    perm = ((k % 3) * 3 + k // 3).astype(f)
    mu = f(0.05) + f(0.20) * perm / f(9)
    sigma = f(0.03) + f(0.12) * perm / f(9)
    weights = np.where(k % 3 == k // 3, f(0.5), f(0.25)).astype(f)
    return (np.stack(spatial), np.stack(spectra), seed, mu.astype(f), sigma.astype(f), weights,
            (k % 3).astype(np.uint32), (k // 3).astype(np.uint32))


def main():
    pkg = ge.load_package()
    cases = [
        # name, n, C, src, dst, radius, temp, patch
        ("c3k9_n16", 16, 3, [0, 1, 2, 0, 1, 2, 0, 1, 2], [0, 0, 0, 1, 1, 1, 2, 2, 2], 5, 1.0, 8),
        ("c2k4_n32", 32, 2, [0, 1, 0, 1], [0, 0, 1, 1], 7, 1.0, 16),
        ("c3k5_n16_ragged", 16, 3, [2, 2, 0, 1, 2], [1, 1, 1, 0, 1], 4, 0.7, 8),
    ]
    for name, n, C, src, dst, R, temp, patch in cases:
        K = len(src)
        cr = pkg.generate_creature(n, C, K, np.array(src), np.array(dst), R, patch=patch, seed=7,
                                   creature=len(name))
        s0 = cr["state"].astype(np.float64)
        s1, u0 = brute_step(s0, cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"], temp)
        s2, _ = brute_step(s1, cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"], temp)
        np.savez_compressed(os.path.join(HERE, f"step_{name}.npz"), state=cr["state"], kernels_fft=cr["kernels_fft"],
                            src=cr["src"], dst=cr["dst"], mu=cr["mu"], sigma=cr["sigma"], weights=cr["weights"],
                            temp=np.float32(temp), conv0=u0, state1=s1, state2=s2)
        print(name, "mass", s0.sum(), s1.sum(), s2.sum())
    # The reference CPU example's own creature at n = 16 and one brute-force step from it.
    spatial, spectra, seed, mu, sigma, w, src, dst = example_creature_brute(16)
    state = seed.astype(np.float32)
    kf = spectra.reshape(9, -1).astype(np.complex64)
    s1, _ = brute_step(state.astype(np.float64), kf, src, dst, mu, sigma, w, 1.0)
    np.savez_compressed(os.path.join(HERE, "example_n16.npz"), spatial=spatial, spectra=spectra.reshape(9, -1),
                        seed=seed, state=state, kernels_fft=kf, mu=mu, sigma=sigma, weights=w, src=src, dst=dst,
                        state1=s1)
    print("example_n16", "mass", state.sum(), s1.sum())


if __name__ == "__main__":
    main()
