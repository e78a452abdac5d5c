"""GPU checks of the device-side creature search: batch-wide parameter copies and cn_adam_step.

- cn_batch_get_param_grads equals the per-creature gradients (bit for bit);
- cn_adam_step equals a float32 restatement of its arithmetic on the host (bit for bit: FP32
  without contraction), and torch.optim.Adam (the reference's optimiser, train/diff_lenia_org.py)
  with the sanitize box of LeniaParams._sanitize (train/utils/leniaparams.py:
  351-360) applied after each step, to 1e-6 of the parameters' scale;
- an optimisation loop run on the device (rollout_grad + adam_step, no host round trip) equals
  the same loop with the update on the host (batch_param_grads, host Adam, batch_set_growth);
- the batch equals one engine per creature.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

LR, B1, B2, EPS = 5e-3, 0.9, 0.999, 1e-8
SPECS = [(64, 3, 8, 1), (128, 2, 12, 2), (64, 4, 6, 3)]  # (n, C, R, seed)


def _batch(cn, specs=SPECS, T=4):
    crs = [cn.baseline_creature(n, num_channels=C, radius=R, seed=s, patch=n // 2) for n, C, R, s in specs]
    tgs = [cn.baseline_creature(n, num_channels=C, radius=R, seed=s + 40, patch=n // 2)["state"]
           for n, C, R, s in specs]
    b = cn.CreatureBatch(crs)
    for i, cr in enumerate(crs):
        b.load(i, cr)
        b.creature_set_target(i, tgs[i])
    return b, crs, tgs


class HostAdam:
    """float32 restatement of cn_adam_step (same operations in the same order, no FMA)."""

    def __init__(self, n):
        f = np.float32
        self.m, self.v = np.zeros(n, f), np.zeros(n, f)
        self.b1t, self.b2t = f(1), f(1)

    def step(self, p, g, K_list):
        f = np.float32
        b1, b2 = f(B1), f(B2)
        self.b1t, self.b2t = f(self.b1t * b1), f(self.b2t * b2)
        self.m = b1 * self.m + (f(1) - b1) * g
        self.v = b2 * self.v + (f(1) - b2) * (g * g)
        mh, vh = self.m / (f(1) - self.b1t), self.v / (f(1) - self.b2t)
        x = p - f(LR) * mh / (np.sqrt(vh) + f(EPS))
        o = 0
        for K in K_list:  # sanitize box per creature: mu in [0, 2], sigma >= 1e-3, w >= 0
            x[o:o + K] = np.minimum(np.maximum(x[o:o + K], f(0)), f(2))
            x[o + K:o + 2 * K] = np.maximum(x[o + K:o + 2 * K], f(1e-3))
            x[o + 2 * K:o + 3 * K] = np.maximum(x[o + 2 * K:o + 3 * K], f(0))
            o += 3 * K
        return x.astype(f)


def _klist(crs):
    return [cr["num_kernels"] for cr in crs]


def test_batch_param_grads_equal_per_creature_grads(cn):
    b, crs, _ = _batch(cn)
    b.rollout_grad(4, 2)
    flat = b.batch_param_grads()
    o = 0
    for i, cr in enumerate(crs):
        K = cr["num_kernels"]
        g = b.creature_grads(i, with_state=False)
        assert np.array_equal(flat[o:o + 3 * K], np.concatenate([g["mu"], g["sigma"], g["weights"]])), i
        o += 3 * K
    b.close()


def test_adam_step_equals_host_restatement_and_torch_adam(cn):
    b, crs, _ = _batch(cn)
    p0 = b.batch_growth()
    host = HostAdam(p0.size)
    tparams = torch.tensor(p0.astype(np.float64), requires_grad=True)
    topt = torch.optim.Adam([tparams], lr=LR, betas=(B1, B2), eps=EPS)
    p = p0
    for it in range(3):
        b.rollout_grad(4, 2)
        g = b.batch_param_grads()
        b.adam_step(LR, B1, B2, EPS)
        got = b.batch_growth()
        p = host.step(p, g, _klist(crs))
        assert np.array_equal(got, p), it
        tparams.grad = torch.tensor(g.astype(np.float64))
        topt.step()
        with torch.no_grad():
            o = 0
            for K in _klist(crs):
                tparams[o:o + K].clamp_(0.0, 2.0)
                tparams[o + K:o + 2 * K].clamp_(min=1e-3)
                tparams[o + 2 * K:o + 3 * K].clamp_(min=0.0)
                o += 3 * K
        ref = tparams.detach().numpy()
        # relative to the parameters' scale: the device keeps them in float32
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"iteration {it}: device Adam vs torch.optim.Adam (float64) max abs / max |p| {err:.2e}")
        assert err <= 1e-6
    b.close()


def test_device_loop_equals_host_loop(cn):
    dev, crs, _ = _batch(cn)
    hst, _, _ = _batch(cn)
    adam = HostAdam(dev.batch_growth().size)
    for it in range(3):
        dev.rollout_grad(4, 2)
        dev.adam_step(LR, B1, B2, EPS)  # no host round trip
        hst.rollout_grad(4, 2)
        p = adam.step(hst.batch_growth(), hst.batch_param_grads(), _klist(crs))
        hst.batch_set_growth(p)
    dev.rollout_grad(4, 2)
    hst.rollout_grad(4, 2)
    assert np.array_equal(dev.batch_growth(), hst.batch_growth())
    for i in range(len(crs)):
        assert dev.creature_loss(i) == hst.creature_loss(i), i
    dev.close()
    hst.close()


def test_batch_adam_loop_equals_single_engines(cn):
    b, crs, tgs = _batch(cn)
    for _ in range(2):
        b.rollout_grad(4, 2)
        b.adam_step(LR, B1, B2, EPS)
    b.rollout_grad(4, 2)
    flat = b.batch_growth()
    o = 0
    for i, cr in enumerate(crs):
        e = cn.CreatureBatch([cr])
        e.load(0, cr)
        e.creature_set_target(0, tgs[i])
        for _ in range(2):
            e.rollout_grad(4, 2)
            e.adam_step(LR, B1, B2, EPS)
        e.rollout_grad(4, 2)
        K = cr["num_kernels"]
        assert np.array_equal(e.batch_growth(), flat[o:o + 3 * K]), i
        assert e.creature_loss(0) == b.creature_loss(i), i
        o += 3 * K
        e.close()
    b.close()


def test_set_growth_after_device_step_keeps_other_creatures(cn):
    """A per-creature upload after a device step must not overwrite the other creatures' device
    parameters with stale host copies."""
    b, crs, _ = _batch(cn)
    b.rollout_grad(4, 2)
    b.adam_step(LR, B1, B2, EPS)
    after = b.batch_growth()
    K0 = crs[0]["num_kernels"]
    b.creature_set_growth(1, crs[1]["mu"], crs[1]["sigma"], crs[1]["weights"])
    now = b.batch_growth()
    assert np.array_equal(now[:3 * K0], after[:3 * K0])
    K1 = crs[1]["num_kernels"]
    assert np.array_equal(now[3 * K0:3 * K0 + K1], crs[1]["mu"])
    b.close()


def test_adam_step_requires_a_backward(cn):
    b, _, _ = _batch(cn)
    with pytest.raises(cn.CellaNovaError) as ei:
        b.adam_step()
    assert ei.value.status == 7
    with pytest.raises(cn.CellaNovaError):
        b.batch_param_grads()
    b.close()
