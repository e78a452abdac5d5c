"""cn_mass_fp64_noalloc: the allocation-free FP64 channel mass of a caller's device state.

Checked against the engine's own cn_mass_fp64 (same per-block partials and the same combination
order, so the sums agree bit for bit) on two calls in a row, the second on a smaller state, so
that nothing of the first call's partials leaks into the second.  Argument checks run without a
GPU."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _call(cn, ptr, n, C, out, stream=None):
    return cn.lib().cn_mass_fp64_noalloc(ctypes.c_void_p(ptr), n, C, ctypes.c_void_p(out), ctypes.c_void_p(stream))


@pytest.mark.parametrize("C,n,status", [(0, 64, 1), (17, 64, 1), (3, 0, 1), (3, 70000, 1)])
def test_noalloc_rejects_bad_sizes_before_any_hip_call(cn, C, n, status):
    assert _call(cn, 0x1000, n, C, 0x2000) == status
    assert len(cn.lib().cn_last_error()) > 0


def test_noalloc_rejects_null_pointers(cn):
    assert _call(cn, 0, 64, 3, 0x2000) == 1
    assert _call(cn, 0x1000, 64, 3, 0) == 1


@pytest.mark.gpu
def test_noalloc_mass_equals_engine_mass_on_two_calls(cn):
    cases = [(256, 3, 4), (64, 2, 9)]  # (n, C, seed): the second call is smaller
    for n, C, seed in cases:
        cr = cn.baseline_creature(n, num_channels=C, radius=max(4, n // 8), seed=seed, patch=n // 2)
        eng = cn.make_engine(cr)
        eng.run_steps(3)
        eng.sync()
        want = eng.mass_fp64()
        ptr, numel = eng.creature_state_ptr(0)
        assert numel == C * n * n
        out = torch.full((C,), float("nan"), dtype=torch.float64, device="cuda:0")
        s = torch.cuda.current_stream()
        assert _call(cn, ptr, n, C, out.data_ptr(), s.cuda_stream) == 0
        got = out.cpu().numpy()
        print(f"n={n} C={C}: noalloc {got}, engine {want}")
        assert np.array_equal(got, want)
        eng.close()


@pytest.mark.gpu
def test_noalloc_mass_sixteen_channels(cn):
    """The largest channel count the scratch holds, against a numpy FP64 sum."""
    rng = np.random.default_rng(5)
    x = rng.random((16, 32, 32), dtype=np.float32)
    d = torch.tensor(x, device="cuda:0")
    out = torch.zeros(16, dtype=torch.float64, device="cuda:0")
    assert _call(cn, d.data_ptr(), 32, 16, out.data_ptr(), torch.cuda.current_stream().cuda_stream) == 0
    got = out.cpu().numpy()
    want = x.astype(np.float64).reshape(16, -1).sum(axis=1)
    assert np.max(np.abs(got - want) / want) <= 1e-14
