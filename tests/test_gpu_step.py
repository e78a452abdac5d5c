"""GPU parity of the DiffusionLenia step (C ABI -> HIP kernels on cuda:0) against the oracle.

Bar: relative L-inf <= 1e-5 per step against the FP64 oracle on teacher-forced states
(both sides start every step from the same state), FP64 mass conserved to 1e-6 relative."""
import glob
import json
import os
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5
GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "step_*.npz")))

MAPS = {
    "c3k9": ([0, 1, 2, 0, 1, 2, 0, 1, 2], [0, 0, 0, 1, 1, 1, 2, 2, 2]),
    "c2k4": ([0, 1, 0, 1], [0, 0, 1, 1]),
    "c1k1": ([0], [0]),
    "ragged_c3k5": ([2, 2, 0, 1, 2], [1, 1, 1, 0, 1]),   # channel 2 receives no kernel
    "c4k2": ([3, 3], [0, 3]),                           # channels 1, 2 are neither source nor destination
}


def _creature(cn, n, maps, R, seed=0, patch=None):
    src, dst = MAPS[maps]
    C = max(max(src), max(dst)) + 1
    return cn.generate_creature(n, C, len(src), np.array(src), np.array(dst), R, patch=patch or max(4, n // 2),
                                seed=seed, creature=n)


def _args(cr):
    return (cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"])


def _teacher_forced(cn, orc, cr, steps, temp=1.0):
    eng = cn.make_engine(cr, temp=temp)
    shape = cr["state"].shape
    errs = []
    for _ in range(steps):
        prev = eng.download_all_channels().reshape(shape).astype(np.float64)
        eng.iterate()
        got = eng.download_all_channels().reshape(shape)
        want = orc.run_f64(prev, *_args(cr), temp=temp)
        errs.append(orc.rel_linf(got, want))
    eng.close()
    return errs


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256])
@pytest.mark.parametrize("maps", list(MAPS))
def test_step_parity_teacher_forced(cn, orc, n, maps):
    R = max(2, n // 5)
    cr = _creature(cn, n, maps, R, seed=n)
    errs = _teacher_forced(cn, orc, cr, steps=3)
    print(f"n={n} {maps}: rel L-inf per step {['%.2e' % e for e in errs]}")
    assert max(errs) <= TOL


def test_step_parity_c5k25(cn, orc):
    cr = cn.baseline_creature(128, num_channels=5, radius=14, seed=3, patch=64)
    errs = _teacher_forced(cn, orc, cr, steps=2)
    assert max(errs) <= TOL


def shift_kernels(cr, ks, dx=3, dy=-2):
    """Translate kernels `ks` by (dx, dy) cells: their spectra gain a phase, so they are no
    longer real and even and take the full complex spectrum path; the rest stay on the compact
    real copy."""
    n = cr["state"].shape[1]
    j = np.arange(n)
    ph = np.exp(-2j * np.pi * (j[:, None] * dy + j[None, :] * dx) / n)
    kf = np.array(cr["kernels_fft"], dtype=np.complex128)
    for k in ks:
        kf[k] = (kf[k].reshape(n, n) * ph).reshape(kf[k].shape)
    cr["kernels_fft"] = kf.astype(np.complex64)
    return cr


def _real_even(spec):
    n = int(round(np.sqrt(spec.size)))
    spec = spec.reshape(n, n)
    h = 0.5 * (spec + np.conj(spec[np.r_[0, len(spec) - 1:0:-1]][:, np.r_[0, len(spec) - 1:0:-1]]))
    tol = 2.0 ** -23 * np.abs(h).max()
    return np.abs(h.imag).max() <= tol and np.abs(h.real - h.real[np.r_[0, len(h) - 1:0:-1]]).max() <= tol


@pytest.mark.parametrize("n", [64, 256])
def test_compact_and_complex_kernel_paths(cn, orc, n):
    cr = _creature(cn, n, "c3k9", max(4, n // 6), seed=7)
    assert all(_real_even(np.asarray(cr["kernels_fft"][k], np.complex128)) for k in range(9))
    shift_kernels(cr, [1, 5, 6])
    assert not _real_even(np.asarray(cr["kernels_fft"][1], np.complex128))
    errs = _teacher_forced(cn, orc, cr, steps=3)
    print(f"n={n} mixed compact/complex kernels: rel L-inf per step {['%.2e' % e for e in errs]}")
    assert max(errs) <= TOL
    eng = cn.make_engine(cr)
    assert eng.info()["compact_kernels"] == 6
    eng.iterate()
    conv = eng.download_conv_buffer().reshape(9, n, n) / (n * n)
    want = orc.conv_f64(cr["state"].astype(np.float64), cr["kernels_fft"], cr["src"])
    assert orc.rel_linf(conv, want) <= 1e-6
    shift_kernels(cr, [1, 5, 6], dx=-3, dy=2)          # back to centred: all compact again
    for k in (1, 5, 6):
        eng.set_kernel(cr["kernels_fft"][k], k)
    assert eng.info()["compact_kernels"] == 9
    eng.close()


def test_conv_buffer_matches_oracle(cn, orc):
    cr = cn.baseline_creature(64, num_channels=3, radius=9, seed=5, patch=32)
    eng = cn.make_engine(cr)
    assert np.all(eng.download_conv_buffer() == 0.0)  # zeros before the first step
    eng.iterate()
    conv = eng.download_conv_buffer().reshape(9, 64, 64) / (64 * 64)  # unnormalised like the reference
    want = orc.conv_f64(cr["state"].astype(np.float64), cr["kernels_fft"], cr["src"])
    assert orc.rel_linf(conv, want) <= 1e-6
    eng.close()


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_fixtures_on_gpu(cn, orc, path):
    z = np.load(path)
    C = z["state"].shape[0]
    n = z["state"].shape[1]
    K = len(z["src"])
    eng = cn.GpuMaceLenia(n, C, K, z["src"], z["dst"], z["mu"], z["sigma"], z["weights"], temp=float(z["temp"]))
    for k in range(K):
        eng.set_kernel(z["kernels_fft"][k], k)
    for c in range(C):
        eng.upload_channel(z["state"][c], c)
    eng.iterate()
    s1 = eng.download_all_channels().reshape(C, n, n)
    assert orc.rel_linf(s1, z["state1"]) <= TOL
    conv = eng.download_conv_buffer().reshape(K, n, n) / (n * n)
    assert orc.rel_linf(conv, z["conv0"]) <= 1e-6
    eng.iterate()
    s2 = eng.download_all_channels().reshape(C, n, n)
    assert orc.rel_linf(s2, z["state2"]) <= 2 * TOL  # free-running over two steps
    eng.close()


def test_temp0_is_box_mean(cn, orc):
    cr = cn.baseline_creature(64, radius=8, seed=1, patch=32)
    eng = cn.make_engine(cr, temp=0.0)
    eng.iterate()
    got = eng.download_all_channels().reshape(3, 64, 64)
    want = orc.box3(cr["state"].astype(np.float64)) / 9.0
    assert orc.rel_linf(got, want) <= 1e-6
    eng.close()


def test_uniform_state_is_fixed_point(cn):
    cr = cn.baseline_creature(128, radius=12, seed=2)
    cr["state"] = np.full_like(cr["state"], 0.25)
    eng = cn.make_engine(cr)
    eng.run_steps(5)
    got = eng.download_all_channels()
    assert np.abs(got - 0.25).max() <= 1e-6
    eng.close()


def test_graph_replay_equals_direct_launch(cn):
    cr = cn.baseline_creature(128, radius=12, seed=4, patch=64)
    a, b = cn.make_engine(cr), cn.make_engine(cr)
    a.run_steps(4)  # HIP-graph replays
    for _ in range(4):
        ms = b.timed_iterate()  # direct launches with events
        assert len(ms) == 4 and all(x >= 0 for x in ms) and ms[3] == 0.0
    assert np.array_equal(a.download_all_channels(), b.download_all_channels())
    assert a.info()["steps_done"] == 4
    a.close()
    b.close()


@pytest.mark.parametrize("n", [128, 2048])
def test_profile_kernels_steps(cn, n):
    """cn_profile_kernels times the passes of 4 * steps direct steps: positive per-launch
    times, and the state of as many plain steps."""
    cr = cn.baseline_creature(n, radius=max(4, n // 80), seed=6, patch=64)
    a, b = cn.make_engine(cr), cn.make_engine(cr)
    ms = a.profile_kernels(7)
    b.run_steps(28)
    assert ms.shape == (4,) and np.all(ms > 0)
    assert a.info()["steps_done"] == 28
    assert np.array_equal(a.download_all_channels(), b.download_all_channels())
    a.run_steps(3)
    b.run_steps(3)
    assert np.array_equal(a.download_all_channels(), b.download_all_channels())
    a.close()
    b.close()


@pytest.mark.parametrize("n,fwd_e", [(128, 0), (1024, 8), (1024, 16)])
def test_multi_step_graph_equals_single_steps(cn, n, fwd_e):
    """run_steps replays graphs of several steps for long runs: the same states as single
    steps, from either ping-pong parity, with eager work between replays, for the default and
    for tuned row geometries."""
    cr = cn.baseline_creature(n, radius=max(4, n // 10), seed=5, patch=64)
    a, b = cn.make_engine(cr), cn.make_engine(cr)
    if fwd_e:
        a.set_plan(n, fwd_e=fwd_e)
        b.set_plan(n, fwd_e=fwd_e)
    a.iterate()                      # odd parity before the multi-step runs
    a.run_steps(21)                  # 8 + 8 + 4 + 1
    mid = a.download_all_channels()  # eager readback between replays
    a.run_steps(16)                  # 8 + 8
    a.run_steps(7)                   # 4 + 2 + 1, from even parity
    a.run_steps(6)                   # 4 + 2, from odd parity
    for _ in range(51):
        b.iterate()
        if b.info()["steps_done"] == 22:
            assert np.array_equal(b.download_all_channels(), mid)
    assert np.array_equal(a.download_all_channels(), b.download_all_channels())
    assert a.info()["steps_done"] == 51
    a.close()
    b.close()


def test_multi_step_graph_mixed_batch(cn):
    """A batch mixing grid sizes and channel counts (size classes 16, 64, 256): multi-step graphs
    and single steps agree bit for bit."""
    crs = [cn.baseline_creature(n, num_channels=C, radius=max(3, n // 10), seed=n + C, patch=min(64, n // 2))
           for n, C in ((16, 2), (64, 3), (256, 2), (16, 3))]

    def run(chained):
        b = cn.CreatureBatch(crs)
        for i, cr in enumerate(crs):
            b.load(i, cr)
        if chained:
            b.run_steps(17)
        else:
            for _ in range(17):
                b.iterate()
        out = [b.creature_download(i) for i in range(len(crs))]
        b.close()
        return out

    for a, s in zip(run(True), run(False)):
        assert np.array_equal(a, s)


def test_mass_conservation_256(cn):
    cr = cn.baseline_creature(256, radius=20, seed=6)
    eng = cn.make_engine(cr)
    m0 = eng.mass_fp64()
    assert abs(m0.sum() - cr["state"].astype(np.float64).sum()) / m0.sum() < 1e-12
    eng.run_steps(200)
    m1 = eng.mass_fp64()
    drift = abs(m1.sum() - m0.sum()) / m0.sum()
    print(f"256^2 C3K9 200 steps: FP64 mass drift {drift:.3e}, per channel {(m1 - m0) / m0}")
    assert drift <= 1e-6
    eng.close()


def test_upload_download_roundtrip(cn):
    cr = cn.baseline_creature(32, radius=4, seed=8, patch=16)
    eng = cn.make_engine(cr)
    x = np.random.default_rng(0).random((3, 32, 32))
    for c in range(3):
        eng.upload_channel(x[c], c)  # f64 -> f32 on upload (mod.rs:220)
    assert np.array_equal(eng.download_all_channels(), x.astype(np.float32).reshape(-1))
    eng.close()


def test_upload_f64_checks_sizes_before_reading(cn):
    """cn_upload_channel_f64 rejects a length or channel that does not match the creature
    before it allocates or reads: a huge length is an argument error, not a host allocation."""
    cr = cn.baseline_creature(32, radius=4, seed=8, patch=16)
    eng = cn.make_engine(cr)
    buf = np.zeros(32 * 32, np.float64)
    p = buf.ctypes.data
    assert cn.lib().cn_upload_channel_f64(eng._h, 0, p, 1 << 62) == 1  # CN_ERR_INVALID_ARG
    assert b"expected 1024" in cn.lib().cn_last_error()
    assert cn.lib().cn_upload_channel_f64(eng._h, 3, p, 32 * 32) == 1
    assert cn.lib().cn_upload_channel_f64(eng._h, 0, p, 32 * 32) == 0
    eng.close()


def test_drop_in_files_json_and_bin(cn, orc, tmp_path):
    """Reference file formats end to end: seed JSON (c0/c1 swapped at load) and a smaller
    kernel .bin spectrally padded to the grid (config.rs:32-125)."""
    n, S = 64, 32
    cr = cn.baseline_creature(S, radius=6, seed=9, patch=16)
    json_c0 = [int(x) for x in cr["dst"]]  # JSON convention: c0 is the output channel
    json_c1 = [int(x) for x in cr["src"]]
    seed = np.zeros((3, n, n))
    seed[:, 16:48, 16:48] = np.random.default_rng(1).random((3, 32, 32))
    cfg = {"seed_size": n, "num_channels": 3, "num_kernels": 9, "seed_channels": seed.reshape(3, -1).tolist(),
           "c0": json_c0, "c1": json_c1, "growth_mu": cr["mu"].tolist(), "growth_sigma": cr["sigma"].tolist(),
           "growth_weights": cr["weights"].tolist(), "global_r": 10.0}
    (tmp_path / "c.json").write_text(json.dumps(cfg))
    with open(tmp_path / "k.bin", "wb") as f:
        for k in range(9):
            f.write(np.ascontiguousarray(cr["kernels_fft"][k]).view(np.float32).astype("<f4").tobytes())
    conf = cn.load_config(tmp_path / "c.json")
    kers = cn.load_kernels(tmp_path / "k.bin", 9, n)
    eng = cn.GpuMaceLenia([n, n], 3, 9, conf.c0, conf.c1, conf.growth_mu, conf.growth_sigma, conf.growth_weights)
    for k in range(9):
        eng.set_kernel(kers[k], k)
    for c in range(3):
        eng.upload_channel(conf.seed_channels[c], c)
    eng.iterate()
    got = eng.download_all_channels().reshape(3, n, n)
    want = orc.run_f64(seed.astype(np.float32), kers, conf.c0, conf.c1, conf.growth_mu, conf.growth_sigma,
                       conf.growth_weights)
    assert orc.rel_linf(got, want) <= TOL
    eng.close()


@pytest.mark.parametrize("n,R", [(1024, 13), (2048, 26)])
def test_full_size_step_vs_fp64(cn, orc, n, R):
    """The benchmark creatures (1024^2 R=13 and 2048^2 R=26, C3K9): teacher-forced steps against
    the numpy FP64 restatement (itself pinned against the C++ FP64 oracle at small sizes)."""
    cr = cn.baseline_creature(n, radius=R, seed=0)
    eng = cn.make_engine(cr)
    eng.run_steps(20)
    errs = []
    for _ in range(2):
        prev = eng.download_all_channels().reshape(3, n, n).astype(np.float64)
        eng.iterate()
        got = eng.download_all_channels().reshape(3, n, n)
        want = orc.step_numpy_f64(prev, *_args(cr))
        errs.append(orc.rel_linf(got, want))
    print(f"{n}^2 C3K9 R={R}: rel L-inf vs FP64 {['%.2e' % e for e in errs]}")
    assert max(errs) <= TOL
    eng.close()


@pytest.mark.parametrize("inv_e", [8, 16])
def test_conv_buffer_2048_row_pairs(cn, orc, inv_e):
    """At 2048^2 the inverse-row pass transforms two rows per complex FFT; its u-writing form
    (the convolution readback and the adjoint's recompute) against the numpy FP64 convolution,
    and the step under both row geometries against FP64."""
    n = 2048
    cr = cn.baseline_creature(n, radius=26, seed=3)
    eng = cn.make_engine(cr)
    eng.set_plan(n, 16, inv_e, 16, 4, 8)
    prev = eng.download_all_channels().reshape(3, n, n).astype(np.float64)
    eng.iterate()
    conv = eng.download_conv_buffer().reshape(9, n, n) / (n * n)
    want = orc.conv_numpy_f64(prev, cr["kernels_fft"], cr["src"])
    err_u = orc.rel_linf(conv, want)
    got = eng.download_all_channels().reshape(3, n, n)
    err_s = orc.rel_linf(got, orc.step_numpy_f64(prev, *_args(cr)))
    print(f"2048^2 C3K9 inv_e={inv_e}: conv readback vs FP64 {err_u:.2e}, step vs FP64 {err_s:.2e}")
    assert err_u <= 1e-6 and err_s <= TOL
    eng.close()


def test_full_size_mass_2048(cn):
    cr = cn.baseline_creature(2048, radius=26, seed=0)
    eng = cn.make_engine(cr)
    m0 = eng.mass_fp64()
    eng.run_steps(300)
    m1 = eng.mass_fp64()
    drift = abs(m1.sum() - m0.sum()) / m0.sum()
    print(f"2048^2 C3K9 300 steps: FP64 mass drift {drift:.3e}")
    assert drift <= 1e-6
    eng.close()


def test_maximum_channels_and_kernels(cn, orc):
    """The largest creature the ABI accepts: C = 16 channels (the reference growth shader's
    array<f32, 16>) and K = 256 kernels (every source x destination pair), against the FP64
    oracle; mass conserved."""
    cr = cn.baseline_creature(32, num_channels=16, radius=5, seed=40, patch=16)
    assert cr["num_kernels"] == 256
    errs = _teacher_forced(cn, orc, cr, steps=2)
    print(f"32^2 C16K256: rel L-inf per step {['%.2e' % e for e in errs]}")
    assert max(errs) <= TOL
    eng = cn.make_engine(cr)
    m0 = eng.mass_fp64()
    eng.run_steps(10)
    assert np.abs(eng.mass_fp64() - m0).max() <= 1e-6 * np.abs(m0).max()
    eng.close()
