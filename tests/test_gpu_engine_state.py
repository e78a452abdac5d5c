"""GPU checks of the engine's call-order rules and of the stream / device-pointer entry points.

- A new launch plan (set_plan, autotune) drops the captured graphs but not the results of the
  last rollout: the backward of an earlier forward and the gradients of an earlier rollout_grad
  read the same slots as a run without the plan change (bitwise).
- Changing the growth parameters or a kernel spectrum after a rollout forward invalidates it:
  the backward refuses (CN_ERR_STATE) instead of recomputing with the new parameters.
- cn_iterate_on_stream (iterate_with_encoder, ml-rs/orchestrator/mod.rs:435-466) orders the
  step between work the caller queued on its own stream before and after it.
- cn_creature_state_ptr (channel_buffer, mod.rs:202-204) is the current state's address.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CN_ERR_STATE = 7


def _creature(cn, n=64, C=3, seed=3):
    return cn.baseline_creature(n, num_channels=C, radius=max(4, n // 6), seed=seed, patch=n // 2)


def _target(cn, n=64, C=3):
    return cn.baseline_creature(n, num_channels=C, radius=max(4, n // 6), seed=77, patch=n // 2)["state"]


def _grads(eng):
    g = eng.creature_grads(0)
    return {k: g[k].copy() for k in ("state0", "mu", "sigma", "weights")}


def test_backward_after_set_plan_matches_run_without_it(cn):
    """The backward recomputes its segments under the plan in force: with the default plan set
    again (graphs dropped, same geometry) the gradients are bitwise those of a run without the
    call; under another geometry they differ by FFT rounding only."""
    cr = _creature(cn)
    gen = torch.Generator(device="cuda:0").manual_seed(1)
    g_final = torch.rand(3 * 64 * 64, dtype=torch.float32, device="cuda:0", generator=gen) - 0.5
    runs = {}
    for plan in (None, (), (8, 8, 8, 2, 8)):
        eng = cn.make_engine(cr)
        eng.rollout_forward(5, 2)
        if plan is not None:
            eng.set_plan(64, *plan)  # drops every captured graph
        eng.rollout_backward([g_final])
        runs[plan] = _grads(eng)
        eng.close()
    for k in runs[None]:
        assert np.array_equal(runs[None][k], runs[()][k]), k
        ref = runs[None][k].astype(np.float64)
        err = np.max(np.abs(runs[(8, 8, 8, 2, 8)][k] - ref)) / np.max(np.abs(ref))
        assert err <= 1e-5, (k, err)


@pytest.mark.parametrize("steps,ckpt", [(3, 2), (5, 5)])
def test_gradients_after_set_plan_match_run_without_it(cn, steps, ckpt):
    """Odd T puts dL/dstate0 in gradient slot 1: the slot must survive the plan change."""
    cr, tgt = _creature(cn), _target(cn)
    runs = {}
    for change_plan in (False, True):
        eng = cn.make_engine(cr)
        eng.creature_set_target(0, tgt)
        eng.rollout_grad(steps, ckpt)
        if change_plan:
            eng.set_plan(64, 8, 8, 8, 2, 8)
        runs[change_plan] = (_grads(eng), eng.creature_loss(0), eng.creature_rollout_final(0))
        eng.close()
    g0, l0, f0 = runs[False]
    g1, l1, f1 = runs[True]
    assert l0 == l1
    assert np.array_equal(f0, f1)
    for k in g0:
        assert np.array_equal(g0[k], g1[k]), k


def test_autotune_keeps_rollout_results(cn):
    cr, tgt = _creature(cn), _target(cn)
    eng = cn.make_engine(cr)
    eng.creature_set_target(0, tgt)
    eng.rollout_grad(4, 2)
    g0 = _grads(eng)
    eng.autotune(1)
    g1 = _grads(eng)
    eng.close()
    for k in g0:
        assert np.array_equal(g0[k], g1[k]), k


@pytest.mark.parametrize("change", ["growth", "kernel", "state"])
def test_backward_refuses_after_a_change_since_the_forward(cn, change):
    cr = _creature(cn)
    eng = cn.make_engine(cr)
    eng.rollout_forward(4, 2)
    if change == "growth":
        eng.set_growth(cr["mu"] * 1.01, cr["sigma"], cr["weights"])
    elif change == "kernel":
        eng.set_kernel(cr["kernels_fft"][1], 1)
    else:
        eng.upload_channel(cr["state"][0], 0)
    g = torch.zeros(3 * 64 * 64, dtype=torch.float32, device="cuda:0")
    with pytest.raises(cn.CellaNovaError) as ei:
        eng.rollout_backward([g])
    assert ei.value.status == CN_ERR_STATE
    with pytest.raises(cn.CellaNovaError):
        eng.creature_rollout_final(0)
    eng.close()


def test_autograd_forward_validates_before_touching_the_engine(cn):
    """A forward that raises on a bad argument leaves the previous forward's backward usable."""
    from cella_nova_amd import autograd
    cr = _creature(cn)
    eng = cn.make_engine(cr)
    dev = "cuda:0"
    s0 = torch.tensor(cr["state"], device=dev).requires_grad_(True)
    mu, sg, w = (torch.tensor(cr[k], device=dev).requires_grad_(True) for k in ("mu", "sigma", "weights"))
    out = autograd.rollout(eng, s0, mu, sg, w, 4, 2)
    with pytest.raises(ValueError):
        autograd.rollout(eng, s0[:1], mu, sg, w, 4, 2)  # wrong state size
    (out ** 2).mean().backward()  # still the first forward's backward
    assert torch.isfinite(s0.grad).all() and torch.isfinite(mu.grad).all()
    # A newer forward supersedes the first one's backward.
    out1 = autograd.rollout(eng, s0.detach().requires_grad_(True), mu, sg, w, 4, 2)
    out2 = autograd.rollout(eng, s0.detach().requires_grad_(True), mu, sg, w, 4, 2)
    with pytest.raises(RuntimeError):
        (out1 ** 2).mean().backward()
    (out2 ** 2).mean().backward()
    eng.close()


class _DevArray:
    """__cuda_array_interface__ view of a device address (no copy)."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f4", "data": (ptr, False), "version": 3,
                                         "strides": None, "stream": None}


def test_state_ptr_is_the_current_state(cn):
    cr = _creature(cn)
    eng = cn.make_engine(cr)
    p0, n = eng.creature_state_ptr(0)
    assert n == 3 * 64 * 64
    view = torch.as_tensor(_DevArray(p0, n), device="cuda:0")
    torch.cuda.synchronize()
    assert np.array_equal(view.cpu().numpy(), eng.download_all_channels())
    eng.iterate()
    eng.sync()
    p1, _ = eng.creature_state_ptr(0)
    assert p1 != p0  # ping-pong: the step wrote the other buffer
    v1 = torch.as_tensor(_DevArray(p1, n), device="cuda:0")
    assert np.array_equal(v1.cpu().numpy(), eng.download_all_channels())
    eng.close()


def test_iterate_on_stream_orders_caller_work(cn):
    cr = _creature(cn, n=128)
    ref = cn.make_engine(cr)
    eng = cn.make_engine(cr)
    steps = 6
    ref.run_steps(steps)
    want = ref.download_all_channels()
    ref.close()
    s = torch.cuda.Stream()
    ptrs = []
    with torch.cuda.stream(s):
        for _ in range(steps):
            eng.iterate_on_stream(s)
        p, n = eng.creature_state_ptr(0)
        ptrs.append(p)
        # Enqueued on the caller's stream after the steps, with no host wait in between.
        got = torch.as_tensor(_DevArray(p, n), device="cuda:0").clone()
    s.synchronize()
    assert np.array_equal(got.cpu().numpy(), want)
    # The engine's own stream is ordered after the caller's steps too.
    eng.iterate()
    ref2 = cn.make_engine(cr)
    ref2.run_steps(steps + 1)
    assert np.array_equal(eng.download_all_channels(), ref2.download_all_channels())
    ref2.close()
    assert eng.info()["steps_done"] == steps + 1
    h = eng.stream_handle()
    assert h != 0
    eng.iterate_on_stream(h)  # the engine's own stream: a plain step
    eng.iterate_on_stream(None)  # the null stream
    eng.sync()
    torch.cuda.synchronize()
    assert eng.info()["steps_done"] == steps + 3
    eng.close()
