"""GPU: the segmented column pass (overlap-save along the columns, k_col_segments) against the
full-length column transforms and the FP64 oracle.

A kernel's column profile h(y, kx) = sum_ky Khat_h(ky, kx) e^{2 pi i ky y / N} has the kernel's
support in y; the engine finds it from the stored spectrum (tail energy within (2^-22)^2 of the
total, in FP64) and convolves each column on segments of 256 rows (2048^2) or 128 rows (1024^2,
512^2) that overlap by twice that radius.  Checked here: the radius against the same criterion computed with numpy; the
convolution results against the full-length path of the same creature (plan col_full) and
against the FP64 oracle; a kernel without compact support (a wide Gaussian) falls back to the
full-length path, and back again; a batch that mixes both kinds of creatures steps each of them
exactly as its own engine does.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 2048


def _hermitian(kf, n):
    K = kf.reshape(n, n).astype(np.complex128)
    Km = np.conj(np.roll(np.flip(K, (0, 1)), 1, (0, 1)))
    return (0.5 * (K + Km) / n**2).astype(np.complex64).astype(np.complex128)


def _numpy_radius(kfs, n, cap=64):
    """The largest support radius over the kernels (None if one has none within `cap`)."""
    out = -1
    j = np.arange(n)
    dist = np.minimum(j, n - j)
    for kf in kfs:
        Kr = _hermitian(kf, n).real[:, : n // 2 + 1]
        h = np.fft.ifft(Kr, axis=0).real * n
        e = h**2
        tot = e.sum()
        r = next((r for r in range(cap + 1) if e[dist > r].sum() <= 2.0**-44 * tot), None)
        if r is None:
            return None
        out = max(out, r)
    return out


def _wide_gaussian(n, sigma=200.0):
    d = np.minimum(np.arange(n), n - np.arange(n)).astype(np.float64)
    k = np.exp(-(d[:, None] ** 2 + d[None, :] ** 2) / (2 * sigma**2))
    return np.fft.fft2(k / k.sum()).astype(np.complex64).reshape(-1)


@pytest.fixture(scope="module")
def creature(cn):
    return cn.baseline_creature(N, num_channels=3, radius=26, seed=0)


# (grid, channels, ring radius, segment length): the 2048^2 step, the 1024^2 rollout and the
# 512^2 C5K25 search shapes
SHAPES = [(2048, 3, 26, 256), (1024, 3, 13, 128), (512, 5, 12, 128)]


@pytest.mark.parametrize("n,C,R0,M", SHAPES)
def test_radius_and_segments_match_numpy(cn, n, C, R0, M):
    cr = creature_of(cn, n, C, R0)
    eng = cn.make_engine(cr)
    info = eng.info()
    eng.close()
    R = _numpy_radius(cr["kernels_fft"], n)
    assert R is not None and R < R0
    assert info["segment_radius"] == R and info["segment_length"] == M
    assert info["segments"] == -(-n // (M - 2 * R))


def creature_of(cn, n, C, R0):
    return cn.baseline_creature(n, num_channels=C, radius=R0, seed=0, patch=min(n // 2, 256))


@pytest.mark.parametrize("n,C,R0,M", SHAPES)
def test_segmented_convolution_matches_full_path_and_oracle(cn, orc, n, C, R0, M):
    cr = creature_of(cn, n, C, R0)
    seg = cn.make_engine(cr)
    full = cn.make_engine(cr)
    full.set_plan(n, col_full=1)
    assert seg.info()["segments"] > 0
    K = cr["num_kernels"]
    seg.iterate()
    full.iterate()
    # the convolution buffer of the step just taken, unnormalised like the reference's
    us = seg.download_conv_buffer().reshape(K, n, n) / n**2
    uf = full.download_conv_buffer().reshape(K, n, n) / n**2
    want_all = orc.conv_f64(cr["state"].astype(np.float64), cr["kernels_fft"], cr["src"])
    for k in range(K):
        want = want_all[k]
        es = orc.rel_linf(us[k], want)
        ef = orc.rel_linf(uf[k], want)
        print(f"kernel {k}: segmented {es:.2e}, full-length {ef:.2e} (rel L-inf vs FP64)")
        assert es <= 2e-6 and ef <= 2e-6
        assert es <= 2 * ef + 2e-7
    # steps: both paths stay within the step bar of each other over a short run
    seg.run_steps(20)
    full.run_steps(20)
    assert orc.rel_linf(seg.download_all_channels(), full.download_all_channels().astype(np.float64)) <= 1e-5
    # and a step from the same state against the FP64 oracle
    prev = seg.download_all_channels().reshape(cr["state"].shape).astype(np.float64)
    seg.iterate()
    want = orc.run_f64(prev, cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"])
    assert orc.rel_linf(seg.download_all_channels().reshape(want.shape), want) <= 1e-5
    seg.close()
    full.close()


def test_non_compact_kernel_falls_back_and_back(cn, creature):
    kfs = creature["kernels_fft"]
    eng = cn.make_engine(creature)
    eng.run_steps(3)  # graphs captured with the segmented pass
    eng.set_kernel(_wide_gaussian(N), 4)
    assert eng.info()["segments"] == 0 and eng.info()["compact_kernels"] == 9
    eng.run_steps(2)
    s_mixed = eng.download_all_channels()
    eng.set_kernel(kfs[4], 4)
    assert eng.info()["segments"] > 0
    eng.run_steps(2)
    got = eng.download_all_channels()
    eng.close()
    # the same sequence on engines built for each phase: bit for bit
    ref = cn.make_engine(creature)
    ref.run_steps(3)
    state3 = ref.download_all_channels()
    ref.close()
    cr_g = dict(creature)
    cr_g["kernels_fft"] = np.array(kfs).copy()
    cr_g["kernels_fft"][4] = _wide_gaussian(N)
    cr_g["state"] = state3.reshape(creature["state"].shape)
    e2 = cn.make_engine(cr_g)
    e2.run_steps(2)
    assert np.array_equal(e2.download_all_channels(), s_mixed)
    e2.close()
    cr_b = dict(creature)
    cr_b["state"] = s_mixed.reshape(creature["state"].shape)
    e3 = cn.make_engine(cr_b)
    e3.run_steps(2)
    assert np.array_equal(e3.download_all_channels(), got)
    e3.close()


def test_mixed_batch_steps_each_creature_as_its_engine(cn, creature):
    other = dict(creature)
    other["kernels_fft"] = np.array(creature["kernels_fft"]).copy()
    other["kernels_fft"][2] = _wide_gaussian(N)
    crs = [creature, other]
    b = cn.CreatureBatch(crs)
    for i, cr in enumerate(crs):
        b.load(i, cr)
    b.run_steps(3)
    for i, cr in enumerate(crs):
        e = cn.make_engine(cr)
        assert (e.info()["segments"] > 0) == (i == 0)
        e.run_steps(3)
        assert np.array_equal(b.creature_download(i), e.download_all_channels()), i
        e.close()
    b.close()


def test_generated_creature_has_segments(cn):
    spec = cn.creature_spec(N, 3, 26, seed=0)
    b = cn.CreatureBatch.generated([spec])
    info = b.info()
    b.close()
    assert info["segments"] > 0 and info["segment_radius"] < 26


def _conv_errors(cn, orc, cr, n):
    eng = cn.make_engine(cr)
    info = eng.info()
    eng.iterate()
    K = cr["num_kernels"]
    u = eng.download_conv_buffer().reshape(K, n, n) / n**2
    eng.close()
    want = orc.conv_f64(cr["state"].astype(np.float64), cr["kernels_fft"], cr["src"])
    return info, [orc.rel_linf(u[k], want[k]) for k in range(K)]


def test_per_kernel_radii(cn, orc):
    """Kernels of radii 4..30 in one creature: each table is truncated at its own support, the
    segments overlap by the largest."""
    n = 1024
    radii = np.array([4, 9, 13, 30], np.uint32)
    cr = cn.generate_creature(n, 2, 4, np.array([0, 1, 0, 1]), np.array([0, 0, 1, 1]), radii, patch=256, seed=3,
                              creature=1)
    info, errs = _conv_errors(cn, orc, cr, n)
    R = _numpy_radius(cr["kernels_fft"], n)
    assert info["segments"] > 0 and info["segment_radius"] == R and R < 30
    assert info["segments"] == -(-n // (info["segment_length"] - 2 * R))
    assert max(errs) <= 2e-6, errs


def test_radius_beyond_the_cap_uses_full_length_transforms(cn, orc):
    """A kernel whose column support exceeds a quarter of the segment length (32 rows at 1024^2)
    keeps the full-length column pass."""
    n = 1024
    cr = cn.generate_creature(n, 1, 1, np.array([0]), np.array([0]), 40, patch=256, seed=4, creature=2)
    info, errs = _conv_errors(cn, orc, cr, n)
    assert info["segments"] == 0
    assert max(errs) <= 2e-6, errs


@pytest.mark.parametrize("case", range(8))
def test_random_creatures_step_against_fp64(cn, case):
    """Seeded random creatures at 512^2 and 1024^2: 1-4 channels, random kernel maps (channels
    that feed no kernel or receive none included), per-kernel radii 3-30 (all within the 32-row
    cap), so that segment counts from 5 to 15 per column and ragged kernel groups occur.  One step
    from the seed against the float64 restatement of the reference step (torch, on the GPU)."""
    torch = pytest.importorskip("torch")
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    try:
        import torch_ref
    finally:
        sys.path.pop(0)
    rng = np.random.default_rng(1000 + case)
    n = int(rng.choice([512, 1024]))
    C = int(rng.integers(1, 5))
    K = int(rng.integers(1, 2 * C + 2))
    src = rng.integers(0, C, K)
    dst = rng.integers(0, C, K)
    radii = rng.integers(3, 31, K).astype(np.uint32)
    cr = cn.generate_creature(n, C, K, src, dst, radii, patch=min(n // 2, 256), seed=case, creature=case)
    eng = cn.make_engine(cr)
    info = eng.info()
    eng.iterate()
    got = eng.download_all_channels().reshape(C, n, n)
    eng.close()
    dev = "cuda:0"
    kf, s_, d_, mu, sg, w = torch_ref.creature_tensors(cr, dev)
    want = torch_ref.step(torch.tensor(cr["state"], dtype=torch.float64, device=dev), kf, s_, d_, mu, sg, w)
    want = want.cpu().numpy()
    err = float(np.abs(got - want).max() / np.abs(want).max())
    print(f"case {case}: {n}^2 C{C} K{K} radii {radii.tolist()} segments {info['segments']} "
          f"(radius {info['segment_radius']}): rel L-inf {err:.2e}")
    assert err <= 1e-5



@pytest.mark.parametrize("radii,segmented", [([62, 40, 8], True), ([70, 12], False)])
def test_2048_radii_at_the_cap(cn, orc, radii, segmented):
    """2048^2 kernels of radius up to 62 (support 60, inside the 64-row cap: 256-row segments
    that overlap by 120 rows, 16 per column) take the segmented pass; a radius-70 kernel keeps
    the full-length pass.  Every convolution result within 2e-6 of the FP64 oracle (measured
    3.7-3.8e-7)."""
    K = len(radii)
    src, dst = np.array([0, 1, 0][:K]), np.array([1, 0, 0][:K])
    cr = cn.generate_creature(N, 2, K, src, dst, np.array(radii, np.uint32), patch=256, seed=9, creature=5)
    info, errs = _conv_errors(cn, orc, cr, N)
    R = _numpy_radius(cr["kernels_fft"], N)
    print(f"radii {radii}: support {R}, segments {info['segments']} (radius {info['segment_radius']}), "
          f"rel L-inf vs FP64 {max(errs):.2e}")
    if segmented:
        assert R is not None and R <= 64 and info["segment_radius"] == R
        assert info["segments"] == -(-N // (256 - 2 * R)) and info["segment_length"] == 256
    else:
        assert R is None and info["segments"] == 0
    assert max(errs) <= 2e-6, errs
