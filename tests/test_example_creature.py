"""The reference CPU example's creature (ml-rs/example.rs:310-491) in the product generator
(cn_generate_example_creature, FP64) and in the oracle's f32 restatement of the example's own
arithmetic, both against the brute-force fixture tests/golden/example_n16.npz (closed-form FP64
kernels, an explicit-index ifftshift, a matrix DFT; tests/golden/make_golden.py), against the
invariants the construction fixes (each kernel sums to 1, is even on the torus so its spectrum
is real), and one brute-force step from the fixture against the FP64 oracle."""
import os

import numpy as np
import pytest

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "example_n16.npz")


@pytest.fixture(scope="module")
def fx():
    return np.load(FIX)


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


def test_product_generator_matches_brute_force_fixture(cn, fx):
    cr = cn.example_creature(16, spatial=True)
    for k in ("mu", "sigma", "weights", "src", "dst"):
        assert np.array_equal(cr[k], fx[k]), k  # the example's own f32 values
    assert _rel(cr["spatial"], fx["spatial"]) <= 1e-13
    assert _rel(cr["kernels_fft"], fx["spectra"]) <= 2 ** -23  # the FP64 transform, rounded once
    assert _rel(cr["state"], fx["seed"]) <= 2 ** -23


def test_oracle_restatement_matches_brute_force_fixture(orc, fx):
    ex = orc.example_creature_f32(16)
    for k in ("mu", "sigma", "weights", "src", "dst"):
        assert np.array_equal(ex[k], fx[k]), k
    assert _rel(ex["kernels_fft"], fx["spectra"]) <= 1e-6  # f32 kernels, sum and radix-2 FFT
    assert _rel(ex["state"], fx["seed"]) <= 1e-6


def test_example_kernels_invariants(cn):
    cr = cn.example_creature(64, spatial=True)
    sp = cr["spatial"]
    assert np.allclose(sp.sum(axis=(1, 2)), 1.0, atol=1e-12)
    # Even on the torus about the origin, so the spectrum is real (and the DC term is the sum).
    assert np.array_equal(sp, np.roll(sp[:, ::-1, :], 1, axis=1))
    assert np.array_equal(sp, np.roll(sp[:, :, ::-1], 1, axis=2))
    assert np.max(np.abs(cr["kernels_fft"].imag)) <= 1e-6
    assert np.allclose(cr["kernels_fft"][:, 0].real, 1.0, atol=1e-6)
    # Ring kernels: the mass sits near half the scaled radius 10 r_k / 2 (r = 0.25 .. 0.85).
    n = 64
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    d = np.sqrt(np.minimum(yy, n - yy) ** 2.0 + np.minimum(xx, n - xx) ** 2.0)
    mean_r = (sp * d).sum(axis=(1, 2))
    assert np.all(np.diff(mean_r) > 0)  # radii grow with k
    assert 0.5 < mean_r[0] < 3.0 and 2.0 < mean_r[-1] < 6.0


def test_oracle_step_on_the_example_creature(orc, fx):
    got = orc.run_f64(fx["state"].astype(np.float64), fx["kernels_fft"], fx["src"], fx["dst"], fx["mu"], fx["sigma"],
                      fx["weights"])
    assert _rel(got, fx["state1"]) <= 1e-12


def test_planted_missing_ifftshift_fails_the_fixture(fx):
    """A generator that forgot to move the kernel centre to the origin (a plausible mistake)
    is caught: the transform of the centred kernel carries a (-1)^(x+y) factor."""
    n, h = 16, 8
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    centred = fx["spatial"][0][(ii + h) % n, (jj + h) % n]  # undo the shift
    wrong = np.fft.fft2(centred).reshape(-1)
    assert _rel(wrong, fx["spectra"][0]) > 0.5
