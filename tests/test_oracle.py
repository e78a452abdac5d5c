"""Pins the CPU oracle (oracle/) without trusting it: brute-force DFT and direct convolution,
closed forms of the growth bump, textbook special cases of the diffusion, the mass identity,
the committed golden fixtures, and planted mistakes that these checks must catch."""
import glob
import os

import numpy as np
import pytest

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "step_*.npz")))


def _creature(cn, n=16, C=3, R=5, seed=1, patch=8, **kw):
    return cn.baseline_creature(n, num_channels=C, radius=R, seed=seed, patch=patch, **kw)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_fft2_f64_matches_direct_dft(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    assert orc.rel_linf(orc.fft2_f64(x), orc.dft2_direct(x)) < 1e-13


@pytest.mark.parametrize("n", [16, 64])
def test_fft2_f32_reference_arithmetic_close(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    assert orc.rel_linf(orc.fft2_f32(x), orc.dft2_direct(x)) < 2e-6


def test_fft_special_cases(orc):
    n = 16
    d = np.zeros((n, n), complex)
    d[0, 0] = 1.0
    assert np.allclose(orc.fft2_f64(d), np.ones((n, n)), atol=1e-15)  # delta -> constant
    y, x = np.mgrid[0:n, 0:n]
    tone = np.exp(2j * np.pi * (3 * y + 5 * x) / n)
    F = orc.fft2_f64(tone)
    want = np.zeros((n, n), complex)
    want[3, 5] = n * n  # single tone -> delta at its frequency
    assert np.abs(F - want).max() < 1e-11


def test_planted_sign_error_is_caught(orc):
    """A forward transform with the inverse sign convention must fail the DFT pin."""
    n = 8
    rng = np.random.default_rng(0)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    wrong = np.conj(orc.fft2_f64(np.conj(x)))  # = DFT with exp(+i...)
    assert orc.rel_linf(wrong, orc.dft2_direct(x)) > 0.1


# ---------------------------------------------------------------- convolution
def test_conv_matches_direct_circular_convolution(cn, orc):
    cr = _creature(cn, spatial=True)
    n, K = 16, 9
    st = np.random.default_rng(3).random((3, n, n))
    u = orc.conv_f64(st, cr["kernels_fft"], cr["src"])
    for k in range(K):
        kern = np.real(orc.idft2_direct(cr["kernels_fft"][k].reshape(n, n).astype(complex)))
        want = orc.conv_direct(st[cr["src"][k]], kern)
        assert orc.rel_linf(u[k], want) < 1e-12


def test_conv_delta_reproduces_kernel(cn, orc):
    """A unit impulse at (y0, x0) convolves to the spatial kernel shifted to (y0, x0)."""
    cr = _creature(cn, spatial=True)
    n = 16
    st = np.zeros((3, n, n))
    st[1, 5, 9] = 1.0
    u = orc.conv_f64(st, cr["kernels_fft"], cr["src"])
    for k in range(9):
        if cr["src"][k] != 1:
            assert np.abs(u[k]).max() < 1e-15
            continue
        want = np.roll(np.roll(cr["spatial"][k], 5, axis=0), 9, axis=1)
        assert np.abs(u[k] - want).max() < 1e-8  # complex64 spectrum rounding


def test_conv_uses_only_real_part_of_kernel(orc):
    """Re(ifft2(fft2(s) K)) depends only on the Hermitian part of K: equals the direct
    convolution with Re(spatial kernel) even for a complex spatial kernel."""
    n = 8
    rng = np.random.default_rng(5)
    kspat = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    kf = orc.dft2_direct(kspat).astype(np.complex64)
    st = rng.random((1, n, n))
    u = orc.conv_f64(st, kf[None], [0])
    kern = np.real(orc.idft2_direct(kf.astype(complex)))
    assert orc.rel_linf(u[0], orc.conv_direct(st[0], kern)) < 1e-12


def test_conv_linearity(cn, orc):
    cr = _creature(cn)
    rng = np.random.default_rng(9)
    a, b = rng.random((3, 16, 16)), rng.random((3, 16, 16))
    ua = orc.conv_f64(a, cr["kernels_fft"], cr["src"])
    ub = orc.conv_f64(b, cr["kernels_fft"], cr["src"])
    uab = orc.conv_f64(2.0 * a - 3.0 * b, cr["kernels_fft"], cr["src"])
    assert np.abs(uab - (2.0 * ua - 3.0 * ub)).max() < 1e-12


# ---------------------------------------------------------------- growth
def test_growth_closed_form(orc):
    n = 4
    mu, sigma, w, temp = np.float32(0.11), np.float32(0.04), np.float32(1.0), 0.5
    m, s = float(mu), float(sigma)
    u = np.array([[m, m + s, m - s, m + 50 * s]] * n)[None]
    A = orc.growth_f64(u, [0], [mu], [sigma], [w], temp, 0)
    g = np.log(A[0]) / temp
    assert abs(g[0] - 1.0) < 1e-12  # G(mu) = 1
    assert abs(g[1] - (2 * np.exp(-0.5) - 1)) < 1e-12  # G(mu +- sigma) = 2e^-1/2 - 1
    assert abs(g[2] - (2 * np.exp(-0.5) - 1)) < 1e-12
    assert abs(g[3] + 1.0) < 1e-12  # far tail -> -1


def test_growth_weighted_sum_by_destination(orc):
    n = 4
    rng = np.random.default_rng(2)
    u = rng.random((3, n, n)) * 0.3
    mu, sg, w = [0.1, 0.12, 0.08], [0.03, 0.05, 0.04], [0.2, 0.5, 0.3]
    dst = [1, 0, 1]
    A1 = orc.growth_f64(u, dst, mu, sg, w, 1.0, 1)

    def G(k):
        m, s = float(np.float32(mu[k])), float(np.float32(sg[k]))
        return 2 * np.exp(-0.5 * (u[k] - m) ** 2 / s ** 2) - 1

    want = np.exp(float(np.float32(w[0])) * G(0) + float(np.float32(w[2])) * G(2))
    assert np.abs(A1 - want).max() < 1e-12
    A2 = orc.growth_f64(u, dst, mu, sg, w, 1.0, 2)  # no kernel writes channel 2
    assert np.all(A2 == 1.0)


# ---------------------------------------------------------------- diffusion
def test_diffusion_temp0_is_box_mean(orc):
    rng = np.random.default_rng(4)
    s = rng.random((12, 12))
    out = orc.diffusion_f64(np.ones((12, 12)), s)
    assert np.abs(out - orc.box3(s) / 9.0).max() < 1e-15


def test_diffusion_uniform_fixed_point(orc):
    rng = np.random.default_rng(4)
    s = np.full((8, 8), 0.37)
    out = orc.diffusion_f64(np.full((8, 8), rng.random() + 0.5), s)
    assert np.abs(out - s).max() < 1e-15


def test_diffusion_mass_identity_and_bruteforce(orc):
    rng = np.random.default_rng(6)
    A = np.exp(rng.standard_normal((16, 16)))
    s = rng.random((16, 16))
    out = orc.diffusion_f64(A, s)
    assert abs(out.sum() - s.sum()) / s.sum() < 1e-13
    Z = np.zeros_like(A)
    for i in range(16):
        for j in range(16):
            Z[i, j] = sum(A[(i + a) % 16, (j + b) % 16] for a in (-1, 0, 1) for b in (-1, 0, 1))
    want = np.zeros_like(A)
    for i in range(16):
        for j in range(16):
            want[i, j] = A[i, j] * sum(s[(i + a) % 16, (j + b) % 16] / Z[(i + a) % 16, (j + b) % 16]
                                       for a in (-1, 0, 1) for b in (-1, 0, 1))
    assert np.abs(out - want).max() < 1e-14


def test_planted_transposed_stencil_is_caught(orc):
    """A diffusion that gathers with Z from the output site instead of the neighbour breaks the
    mass identity; the identity test would catch it."""
    rng = np.random.default_rng(6)
    A = np.exp(rng.standard_normal((16, 16)))
    s = rng.random((16, 16))
    wrong = A * orc.box3(s) / orc.box3(A)
    assert abs(wrong.sum() - s.sum()) / s.sum() > 1e-3


# ---------------------------------------------------------------- full step
@pytest.mark.parametrize("C,n,R", [(3, 16, 5), (2, 32, 7), (1, 16, 4)])
def test_step_f64_matches_numpy_restatement(cn, orc, C, n, R):
    cr = cn.baseline_creature(n, num_channels=C, radius=R, seed=11, patch=n // 2)
    args = (cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"])
    st = cr["state"].astype(np.float64)
    a = orc.run_f64(st, *args, steps=1)
    b = orc.step_numpy_f64(st, *args)
    assert orc.rel_linf(a, b) < 1e-12
    assert abs(a.sum() - st.sum()) / st.sum() < 1e-13


def test_step_f32_reference_arithmetic_close_to_f64(cn, orc):
    cr = _creature(cn, n=32, R=7, patch=16)
    args = (cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"])
    a = orc.run_f32(cr["state"], *args)
    b = orc.run_f64(cr["state"], *args)
    assert orc.rel_linf(a, b) < 1e-5


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_fixtures(orc, path):
    z = np.load(path)
    args = (z["kernels_fft"], z["src"], z["dst"], z["mu"], z["sigma"], z["weights"])
    s0 = z["state"].astype(np.float64)
    u = orc.conv_f64(s0, z["kernels_fft"], z["src"])
    assert orc.rel_linf(u, z["conv0"]) < 1e-12
    s1 = orc.run_f64(s0, *args, temp=float(z["temp"]), steps=1)
    assert orc.rel_linf(s1, z["state1"]) < 1e-12
    s2 = orc.run_f64(s0, *args, temp=float(z["temp"]), steps=2)
    assert orc.rel_linf(s2, z["state2"]) < 1e-12
    f1 = orc.run_f32(z["state"], *args, temp=float(z["temp"]), steps=1)
    assert orc.rel_linf(f1, z["state1"]) < 1e-5


def test_golden_fixture_planted_wrong_mapping_is_caught(orc):
    """Swapping the src/dst maps (the reference's c0/c1 convention trap) must fail the fixture."""
    z = np.load(GOLDEN[0])
    s1 = orc.run_f64(z["state"].astype(np.float64), z["kernels_fft"], z["dst"], z["src"], z["mu"], z["sigma"],
                     z["weights"], temp=float(z["temp"]))
    assert orc.rel_linf(s1, z["state1"]) > 1e-3


# --- render epilogue colour pack (render.rs:145-153, compute_1024.wgsl:429-443) ---------------
def test_render_rgb_closed_forms(orc):
    st = np.zeros((3, 2, 2), np.float32)
    st[:, 0, 0] = (0.5, 0.0, 2.0 / 3.0)     # floor(255 sqrt(0.75)) = 220; 1.5 * 2/3 clamps to 1 -> 255
    st[:, 0, 1] = (-1.0, 7.0, np.nan)       # clamp below and above; NaN gives 0 here
    st[:, 1, 0] = (1.0 / 6.0, 0.01, 0.1)    # 255 sqrt(0.25) = 127.5 truncates to 127
    rgb = orc.render_rgb(st)
    assert rgb[0, 0].tolist() == [220, 0, 255]
    assert rgb[0, 1].tolist() == [0, 255, 0]
    assert rgb[1, 0].tolist() == [127, int(np.floor(255 * np.sqrt(0.015))), int(np.floor(255 * np.sqrt(0.15)))]
    assert rgb[1, 1].tolist() == [0, 0, 0]


def test_render_rgb_matches_fp64_formula_away_from_integer_boundaries(orc):
    c = np.linspace(-0.05, 0.7, 3 * 40000, dtype=np.float32).reshape(3, 200, 200)
    got = orc.render_rgb(c).transpose(2, 0, 1).astype(np.int64)
    v = 255.0 * np.sqrt(np.clip(1.5 * c.astype(np.float64), 0.0, 1.0))
    want = np.floor(v).astype(np.int64)
    clamped = (c <= 0) | (1.5 * c.astype(np.float64) >= 1.0)
    away = (np.abs(v - np.round(v)) > 1e-4) | clamped   # f32 rounding only matters next to an integer
    assert away.mean() > 0.99 and clamped.mean() < 0.2
    assert np.array_equal(got[away], want[away])
    assert np.abs(got - want).max() <= 1


def test_render_rgb_fewer_channels_and_order(orc):
    st = np.full((1, 4, 4), 0.5, np.float32)
    rgb = orc.render_rgb(st)
    assert (rgb[..., 0] == 220).all() and (rgb[..., 1:] == 0).all()
    st5 = np.stack([np.full((4, 4), v, np.float32) for v in (0.0, 0.5, 2.0 / 3.0, 1.0, 1.0)])
    assert orc.render_rgb(st5)[0, 0].tolist() == [0, 220, 255]   # channels 3.. are not drawn


# --- the PyTorch restatement (oracle/torch_ref.py) against the C++ FP64 oracle -----------------
@pytest.mark.parametrize("C,n", [(1, 32), (2, 32), (3, 64)])
def test_torch_restatement_matches_fp64_oracle(cn, orc, C, n):
    torch = pytest.importorskip("torch")
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import torch_ref
    cr = cn.baseline_creature(n, num_channels=C, radius=max(3, n // 6), seed=C, patch=n // 2)
    want = orc.run_f64(cr["state"].astype(np.float64), cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"],
                       cr["sigma"], cr["weights"], temp=0.7)
    got = torch_ref.step(torch.tensor(cr["state"], dtype=torch.float64), *torch_ref.creature_tensors(cr, "cpu"),
                         temp=float(np.float32(0.7))).numpy()  # the C ABI takes temp as f32
    assert orc.rel_linf(got, want) <= 1e-12
