"""GPU checks of the device creature generator (cn_creature_generate) and the spectrum readback
(cn_creature_download_kernels) against the host FP64 generator and the oracle.

Bars: spectra within 1e-6 of max|K^| of the FP64 spectra rounded to complex64 (f32 FFT rounding);
seed state bit-identical; one step teacher-forced against the FP64 oracle with the host spectra
<= 1e-5 relative L-inf; every generated kernel on the compact real path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [  # n, C, radius (scalar or per kernel), patch
    (64, 3, 9, 32),
    (256, 3, 26, 128),
    (1024, 3, 13, 128),
    (128, 5, [10, 11, 12, 13, 14, 15, 16, 17, 18, 10, 11, 12, 13, 14, 15, 16, 17, 18, 10, 11, 12, 13, 14, 15, 16], 64),
    (64, 2, 40, 64),        # R > N/2: the kernel wraps around the torus
    (16, 1, 3, 8),
]


def _placeholder_engine(cn, n, C, src, dst):
    K = len(src)
    return cn.GpuMaceLenia(n, C, K, src, dst, np.full(K, 0.1, np.float32), np.full(K, 0.05, np.float32),
                           np.full(K, 1.0, np.float32))


@pytest.mark.parametrize("n,C,R,patch", CASES, ids=[f"n{c[0]}_c{c[1]}" + ("_wrap" if c[0] == 64 and c[1] == 2 else "")
                                                    for c in CASES])
def test_device_generator_matches_host(cn, orc, n, C, R, patch):
    src, dst = cn.channel_maps(C)
    cr = cn.generate_creature(n, C, len(src), src, dst, R, patch=patch, seed=5, creature=n + C)
    eng = _placeholder_engine(cn, n, C, src, dst)
    eng.creature_generate(0, R, patch=patch, seed=5, creature=n + C)
    K = len(src)
    assert eng.info()["compact_kernels"] == K
    spec = eng.creature_download_kernels(0)
    ref = np.asarray(cr["kernels_fft"])
    err = np.abs(spec.astype(np.complex128) - ref).max() / np.abs(ref).max()
    print(f"n={n} C={C}: spectra max err {err:.2e} of max|K|")
    assert err <= 1e-6
    st = eng.download_all_channels().reshape(C, n, n)
    assert np.array_equal(st, cr["state"])
    # The growth parameters and spectra in use: one step against the FP64 oracle on host data.
    eng.iterate()
    got = eng.download_all_channels().reshape(C, n, n)
    want = orc.run_f64(cr["state"].astype(np.float64), cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"],
                       cr["sigma"], cr["weights"])
    e = orc.rel_linf(got, want)
    print(f"  step vs FP64 oracle {e:.2e}")
    assert e <= 1e-5
    eng.close()


def test_generated_batch_matches_host_loaded_batch(cn):
    specs = [cn.creature_spec(n, C, R, patch=n // 2, seed=2, creature=i)
             for i, (n, C, R) in enumerate([(64, 3, 8), (128, 2, 12), (256, 4, 20), (64, 3, 9)])]
    crs = [cn.generate_creature(s["n"], s["num_channels"], s["num_kernels"], s["src"], s["dst"], s["radius"],
                                patch=s["patch"], seed=s["seed"], creature=s["creature"]) for s in specs]
    a = cn.CreatureBatch.generated(specs)
    b = cn.CreatureBatch(crs)
    for i, cr in enumerate(crs):
        b.load(i, cr)
        assert np.array_equal(a.creature_download(i), b.creature_download(i))
    a.run_steps(3)
    b.run_steps(3)
    for i in range(len(specs)):
        x, y = a.creature_download(i), b.creature_download(i)
        assert np.abs(x - y).max() <= 1e-5 * np.abs(y).max(), i
    a.close()
    b.close()


def test_download_kernels_returns_hermitian_part_of_set_kernel(cn):
    n, K = 32, 2
    rng = np.random.default_rng(9)
    eng = cn.GpuMaceLenia(n, 1, K, [0, 0], [0, 0], [0.1, 0.1], [0.05, 0.05], [0.5, 0.5])
    full = (rng.standard_normal((K, n, n)) + 1j * rng.standard_normal((K, n, n))).astype(np.complex64)
    for k in range(K):
        eng.set_kernel(full[k].reshape(-1), k)
    got = eng.creature_download_kernels(0).reshape(K, n, n).astype(np.complex128)
    f = full.astype(np.complex128)
    flip = f[:, np.r_[0, n - 1:0:-1]][:, :, np.r_[0, n - 1:0:-1]]
    herm = 0.5 * (f + np.conj(flip))
    assert np.abs(got - herm).max() <= 1e-6 * np.abs(herm).max()
    assert eng.info()["compact_kernels"] == 0
    eng.close()
