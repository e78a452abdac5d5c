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


def test_set_kernel_survives_each_injected_failure(cn):
    """set_kernel at 512^2 also measures the kernel's column support and builds its segment table
    (temporary buffers, three launches, a read-back): each checked call fails in turn, and a
    retry leaves the engine as one that never failed — the same segments, device bytes and
    step results."""
    n = 512
    cr = cn.baseline_creature(n, num_channels=2, radius=9, seed=35, patch=64)
    kf = np.array(cr["kernels_fft"])
    ref = cn.make_engine(cr)
    ref.set_kernel(kf[1], 2)
    want_info = ref.info()
    ref.run_steps(3)
    want = ref.download_all_channels()
    ref.close()
    assert want_info["segments"] > 0

    def same(e):
        info = e.info()
        for key in ("segments", "segment_radius", "device_bytes"):
            assert info[key] == want_info[key], key
        e.run_steps(3)
        assert np.array_equal(e.download_all_channels(), want)

    buf = np.ascontiguousarray(kf[1].astype(np.complex64))
    n_fail = _walk(cn, lambda: cn.make_engine(cr),
                   lambda e: cn.lib().cn_set_kernel(e._h, 2, buf.ctypes.data, buf.size), same)
    assert n_fail >= 8  # the uploads, temporaries, profile, sums, read-back and table at least


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


def test_mixed_batch_creation_and_step_survive_each_injected_failure(cn):
    """A batch of three grid sizes (branch streams and events per size class): every checked call
    of its creation fails in turn and leaves no engine behind; every checked call of its first
    step (graph capture with one branch per class) fails in turn and the same batch then steps to
    the results of a batch that never failed."""
    from cella_nova_amd._lib import CellaNovaError
    crs = [cn.baseline_creature(n, num_channels=C, radius=4, seed=40 + n, patch=min(32, n // 2))
           for n, C in ((16, 2), (32, 3), (64, 2))]

    def make():
        b = cn.CreatureBatch(crs)
        for i, cr in enumerate(crs):
            b.load(i, cr)
        return b

    ref = make()
    ref.iterate()
    want = [ref.creature_download(i) for i in range(len(crs))]
    ref.close()
    lib = cn.lib()
    fails = 0
    for skip in range(200):
        lib.cn_debug_fail_hip_at(skip)
        try:
            b = cn.CreatureBatch(crs)
        except CellaNovaError as err:
            lib.cn_debug_fail_hip_at(-1)
            assert err.status == CN_ERR_HIP and "injected" in err.message, skip
            fails += 1
            continue
        lib.cn_debug_fail_hip_at(-1)
        b.close()
        break
    else:
        pytest.fail("batch creation still reached the injection after 200 checked calls")
    assert fails >= 8  # device, stream, events, descriptor tables, branch streams and events

    def same(b):
        for i in range(len(crs)):
            assert np.array_equal(b.creature_download(i), want[i]), i

    n = _walk(cn, make, lambda b: lib.cn_iterate(b._h), same)
    assert n >= 2
