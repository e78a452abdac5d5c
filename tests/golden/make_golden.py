"""Writes the committed golden fixtures tests/golden/step_*.npz.

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


if __name__ == "__main__":
    main()
