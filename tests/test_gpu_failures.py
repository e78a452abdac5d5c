"""Error paths on the GPU: every checked HIP call and device allocation of a first step and of a
first rollout fails in turn (injected on the host: the failing call is not issued), the entry
point reports CN_ERR_HIP, and the same engine then completes the call with results identical to
an engine that never failed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CN_ERR_HIP = 3


def _engine(cn, cr, target):
    e = cn.make_engine(cr)
    e.creature_set_target(0, target)
    return e


def _walk(cn, make, call, check_same, limit=400):
    """Fails the skip-th checked call of `call` for skip = 0, 1, ... until the call completes
    without reaching the injection; after each failure the same engine must complete it."""
    lib = cn.lib()
    failures = 0
    for skip in range(limit):
        e = make()
        lib.cn_debug_fail_hip_at(skip)
        st = call(e)
        lib.cn_debug_fail_hip_at(-1)
        if st == 0:
            check_same(e)
            e.close()
            return failures
        assert st == CN_ERR_HIP, (skip, st, lib.cn_last_error())
        assert b"injected" in lib.cn_last_error(), skip
        failures += 1
        assert call(e) == 0, (skip, lib.cn_last_error())
        check_same(e)
        e.close()
    pytest.fail(f"the call still reached the injection after {limit} checked calls")


def test_step_survives_each_injected_failure(cn):
    cr = cn.baseline_creature(32, radius=4, seed=31, patch=16)
    ref = cn.make_engine(cr)
    ref.iterate()
    want = ref.download_all_channels()
    ref.close()

    def same(e):
        assert np.array_equal(e.download_all_channels(), want)

    n = _walk(cn, lambda: cn.make_engine(cr), lambda e: cn.lib().cn_iterate(e._h), same)
    assert n >= 2  # hipSetDevice and the capture at least


def test_render_survives_each_injected_failure(cn):
    cr = cn.baseline_creature(32, radius=4, seed=34, patch=16)
    ref = cn.make_engine(cr)
    ref.iterate_and_render()
    want = (ref.download_all_channels(), ref.download_render(), ref.info()["device_bytes"])
    ref.close()

    def same(e):
        assert np.array_equal(e.download_all_channels(), want[0])
        assert np.array_equal(e.download_render(), want[1])
        assert e.info()["device_bytes"] == want[2]

    n = _walk(cn, lambda: cn.make_engine(cr), lambda e: cn.lib().cn_iterate_and_render(e._h), same)
    assert n >= 3  # the frame buffer, its descriptor upload and the capture at least


def test_rollout_survives_each_injected_failure(cn):
    cr = cn.baseline_creature(32, radius=4, seed=32, patch=16)
    target = cn.baseline_creature(32, radius=4, seed=33, patch=16)["state"]
    ref = _engine(cn, cr, target)
    ref.rollout_grad(6, 2)
    want = ref.creature_grads(0)
    want_loss = ref.creature_loss(0)
    want_bytes = ref.info()["device_bytes"]
    ref.close()

    def same(e):
        g = e.creature_grads(0)
        for k in want:
            assert np.array_equal(g[k], want[k]), k
        assert e.creature_loss(0) == want_loss
        # Buffers of a failed attempt are neither kept twice nor counted twice.
        assert e.info()["device_bytes"] == want_bytes

    n = _walk(cn, lambda: _engine(cn, cr, target), lambda e: cn.lib().cn_rollout_grad(e._h, 6, 2), same)
    assert n >= 10  # state and intermediate slots, two captures, the launch
