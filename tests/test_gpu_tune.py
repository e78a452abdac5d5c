"""GPU checks of launch plans and the autotuner: every valid geometry gives parity with the FP64
oracle; the autotuner's choice is valid; invalid plans are rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _args(cr):
    return (cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"])


@pytest.mark.parametrize("n", [64, 256])
def test_every_geometry_keeps_parity(cn, orc, n):
    cr = cn.baseline_creature(n, num_channels=3, radius=max(4, n // 10), seed=n, patch=n // 2)
    want = orc.run_f64(cr["state"].astype(np.float64), *_args(cr))
    ref = None
    for fe in (8, 16):
        for ie in (8, 16):
            for ce, tc in ((8, 2), (8, 4), (16, 4), (16, 8), (16, 16)):
                for th in (8, 16, 32):
                    if tc > n // 2:
                        continue
                    eng = cn.make_engine(cr)
                    eng.set_plan(n, fwd_e=fe, inv_e=ie, col_e=ce, col_tc=tc, diff_th=th)
                    eng.iterate()
                    got = eng.download_all_channels().reshape(cr["state"].shape)
                    eng.close()
                    err = orc.rel_linf(got, want)
                    assert err <= 1e-5, (fe, ie, ce, tc, th, err)
                    if th == 8:
                        # per-(tile, kernel) column blocks do the same arithmetic per column
                        eng = cn.make_engine(cr)
                        eng.set_plan(n, fwd_e=fe, inv_e=ie, col_e=ce, col_tc=tc, diff_th=th, col_split=1)
                        eng.iterate()
                        split = eng.download_all_channels().reshape(cr["state"].shape)
                        eng.close()
                        assert np.array_equal(split, got), (fe, ie, ce, tc)


def test_autotune_choice_is_valid_and_correct(cn, orc):
    cr = cn.baseline_creature(512, num_channels=3, radius=20, seed=3)
    eng = cn.make_engine(cr)
    plan = eng.autotune(3)[0]
    assert plan["n"] == 512 and plan["fwd_e"] in (8, 16) and plan["col_tc"] in (2, 4, 8, 16)
    assert np.array_equal(eng.download_all_channels(), cr["state"].reshape(-1))  # state untouched
    eng.iterate()
    got = eng.download_all_channels().reshape(cr["state"].shape)
    want = orc.step_numpy_f64(cr["state"].astype(np.float64), *_args(cr))
    assert orc.rel_linf(got, want) <= 1e-5
    eng.close()


def test_invalid_plans_rejected(cn):
    cr = cn.baseline_creature(64, num_channels=2, radius=6, seed=1, patch=32)
    eng = cn.make_engine(cr)
    for kw in (dict(fwd_e=4), dict(col_e=32), dict(diff_th=64), dict(col_e=8, col_tc=64), dict(col_split=2)):
        with pytest.raises(cn.CellaNovaError):
            eng.set_plan(64, **kw)
    with pytest.raises(cn.CellaNovaError):
        eng.set_plan(128)  # no creature of that size
    eng.close()


def test_batch_autotune_per_size_class(cn):
    crs = [cn.baseline_creature(n, num_channels=2, radius=6, seed=i, patch=32) for i, n in enumerate((64, 128, 64))]
    b = cn.CreatureBatch(crs)
    for i, cr in enumerate(crs):
        b.load(i, cr)
    plans = b.autotune(2)
    assert sorted(p["n"] for p in plans) == [64, 128]
    b.run_steps(2)
    for i, cr in enumerate(crs):
        e = cn.make_engine(cr)
        for p in plans:
            if p["n"] == cr["n"]:
                e.set_plan(**p)
        e.run_steps(2)
        assert np.array_equal(b.creature_download(i), e.download_all_channels())
        e.close()
    b.close()


def test_split_column_pass_matches_fused_in_mixed_batch(cn):
    """Per-(tile, kernel) column blocks give the per-source pass bit for bit in a batch of
    creatures with different C and K, including kernels on the complex spectrum path."""
    from test_gpu_step import shift_kernels
    crs = [cn.baseline_creature(64, num_channels=C, radius=7, seed=10 + C, patch=32) for C in (2, 3, 4)]
    shift_kernels(crs[1], [1, 4])

    def run(split):
        b = cn.CreatureBatch(crs)
        for i, cr in enumerate(crs):
            b.load(i, cr)
        b.set_plan(64, col_e=16, col_tc=4, col_split=split)
        b.run_steps(3)
        out = [b.creature_download(i) for i in range(len(crs))]
        b.close()
        return out

    for a, s in zip(run(0), run(1)):
        assert np.array_equal(a, s)


@pytest.mark.parametrize("maps", ["c1k1", "ragged_c3k5", "c4k2", "c2k4"])
def test_split_column_pass_matches_fused_for_ragged_maps(cn, maps):
    """Per-kernel column blocks against per-source ones for channel maps with unused sources,
    destinations without kernels and several kernels per (source, destination) pair."""
    from test_gpu_step import MAPS
    src, dst = MAPS[maps]
    C = max(max(src), max(dst)) + 1
    cr = cn.generate_creature(64, C, len(src), np.array(src), np.array(dst), 9, patch=32, seed=7, creature=64)

    def run(split):
        eng = cn.make_engine(cr)
        eng.set_plan(64, col_e=16, col_tc=4, col_split=split)
        eng.run_steps(3)
        out = eng.download_all_channels()
        eng.close()
        return out

    assert np.array_equal(run(0), run(1))
