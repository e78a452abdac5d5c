"""GPU checks of batched execution: creatures of mixed N, C, K advanced in one graph must match
the same creatures run one engine at a time, bit for bit, forward and adjoint."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [  # (n, C, R, seed)
    (64, 3, 8, 1), (128, 2, 12, 2), (64, 4, 6, 3), (256, 3, 20, 4), (128, 5, 10, 5), (64, 3, 9, 6),
]


def _creatures(cn):
    return [cn.baseline_creature(n, num_channels=C, radius=R, seed=s, patch=n // 2) for n, C, R, s in SPECS]


def test_batch_forward_equals_single_engines(cn):
    crs = _creatures(cn)
    b = cn.CreatureBatch(crs)
    for i, cr in enumerate(crs):
        b.load(i, cr)
    b.run_steps(4)
    for i, cr in enumerate(crs):
        e = cn.make_engine(cr)
        e.run_steps(4)
        assert np.array_equal(b.creature_download(i), e.download_all_channels()), i
        m = b.creature_mass(i)
        assert np.allclose(m, e.mass_fp64(), rtol=0, atol=0)
        e.close()
    b.close()


def test_batch_rollout_grad_equals_single_engines(cn):
    crs = _creatures(cn)
    targets = [cn.baseline_creature(n, num_channels=C, radius=R, seed=s + 50, patch=n // 2)["state"]
               for n, C, R, s in SPECS]
    b = cn.CreatureBatch(crs)
    for i, cr in enumerate(crs):
        b.load(i, cr)
        b.creature_set_target(i, targets[i])
    b.rollout_grad(6, 3)
    for i, cr in enumerate(crs):
        e = cn.make_engine(cr)
        e.creature_set_target(0, targets[i])
        e.rollout_grad(6, 3)
        gb, ge = b.creature_grads(i), e.creature_grads(0)
        for k in ("state0", "mu", "sigma", "weights"):
            assert np.array_equal(gb[k], ge[k]), (i, k)
        assert b.creature_loss(i) == e.creature_loss(0)
        e.close()
    # a replay of the captured rollout graph gives the same result
    g0 = b.creature_grads(0)
    b.rollout_grad(6, 3)
    assert np.array_equal(b.creature_grads(0)["state0"], g0["state0"])
    b.close()


def test_batch_set_growth_changes_only_its_creature(cn):
    crs = _creatures(cn)[:3]
    b = cn.CreatureBatch(crs)
    for i, cr in enumerate(crs):
        b.load(i, cr)
    b.creature_set_growth(1, crs[1]["mu"] * 1.1, crs[1]["sigma"], crs[1]["weights"])
    b.run_steps(2)
    e0 = cn.make_engine(crs[0])
    e0.run_steps(2)
    assert np.array_equal(b.creature_download(0), e0.download_all_channels())
    cr1 = dict(crs[1])
    cr1["mu"] = crs[1]["mu"] * 1.1
    e1 = cn.make_engine(cr1)
    e1.run_steps(2)
    assert np.array_equal(b.creature_download(1), e1.download_all_channels())
    e0.close()
    e1.close()
    b.close()


def test_batch_conv_readback_is_creature_0(cn):
    """cn_download_conv_real on a batch recomputes creature 0's convolution alone (the other
    creatures of its size class have no convolution buffer): the same values as a single engine
    of creature 0, and the other creatures' states are untouched."""
    crs = [cn.baseline_creature(64, num_channels=3, radius=8, seed=s, patch=32) for s in (11, 12, 13)]
    b = cn.CreatureBatch(crs)
    for i, cr in enumerate(crs):
        b.load(i, cr)
    b.iterate()
    before = [b.creature_download(i) for i in range(3)]
    got = np.empty(9 * 64 * 64, np.float32)
    assert cn.lib().cn_download_conv_real(b._h, got.ctypes.data, got.size) == 0, cn.lib().cn_last_error()
    e = cn.make_engine(crs[0])
    e.iterate()
    assert np.array_equal(got, e.download_conv_buffer())
    e.close()
    for i in range(3):
        assert np.array_equal(b.creature_download(i), before[i])
    b.close()
