"""The batched benchmark configurations on the GPU against the FP64 oracle.

Uneven batch (256 creatures: 85 at 256^2, 85 at 512^2, 86 at 1024^2; C ~ U{2,3,4}, K = C^2,
R ~ U{8..20}; creatures generated on the device): one teacher-forced step of two creatures of
every (N, C) cell against the FP64 oracle (the C++ restatement of ml-rs-example in double at
256^2, the numpy FP64 restatement of the reference step above), the spectra the creature
steps with read back from the device; FP64 mass of all 256 creatures conserved over 200 steps.

Batched search (32 creatures at 512^2, C5K25, R ~ U{10..18}, T = 32, checkpoint every 4):
gradients of two creatures against torch autograd of the float64 restatement of the
reference model (oracle/torch_ref.py, train/diff_lenia_org.py:27-53), and one device-side Adam
iteration of the batch against single-creature engines, bit for bit.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import torch_ref  # noqa: E402

DEV = "cuda:0"


def _host_params(cn, sp):
    """Growth parameters of a generated creature (the host generator draws the same values)."""
    return cn.generate_creature(sp["n"], sp["num_channels"], sp["num_kernels"], sp["src"], sp["dst"], sp["radius"],
                                patch=sp["patch"], seed=sp["seed"], creature=sp["creature"], spectra=False, state=False)


def test_config3_uneven_batch_against_fp64_oracle(cn, orc):
    specs = cn.creatures.uneven_batch_specs(0)
    assert len(specs) == 256 and [sp["n"] for sp in specs].count(1024) == 86
    b = cn.CreatureBatch.generated(specs)
    cells = {}
    for i, sp in enumerate(specs):
        cells.setdefault((sp["n"], sp["num_channels"]), []).append(i)
    assert len(cells) == 9
    picks = sorted({j for idx in cells.values() for j in (idx[0], idx[-1])})
    m0 = np.array([b.creature_mass(i).sum() for i in range(len(specs))])
    prev = {i: b.creature_download(i) for i in picks}
    b.iterate()
    worst = 0.0
    for i in picks:
        sp = specs[i]
        n, C = sp["n"], sp["num_channels"]
        got = b.creature_download(i).reshape(C, n, n)
        kf = b.creature_download_kernels(i)
        gp = _host_params(cn, sp)
        args = (kf, sp["src"], sp["dst"], gp["mu"], gp["sigma"], gp["weights"])
        x = prev[i].reshape(C, n, n).astype(np.float64)
        want = orc.run_f64(x, *args) if n <= 256 else orc.step_numpy_f64(x, *args)
        err = orc.rel_linf(got, want)
        worst = max(worst, err)
        print(f"creature {i}: {n}^2 C{C}K{C * C} R={sp['radius'][0]}: teacher-forced step vs FP64 {err:.2e}")
        assert err <= 1e-5, (i, err)
    b.run_steps(199)  # 200 steps in all
    m1 = np.array([b.creature_mass(i).sum() for i in range(len(specs))])
    drift = np.abs(m1 - m0) / m0
    print(f"uneven batch: worst per-step error {worst:.2e}; FP64 mass drift over 200 steps max {drift.max():.2e} "
          f"(creature {int(drift.argmax())})")
    assert drift.max() <= 1e-6
    b.close()


def _torch_grads(cr, target, steps):
    kf, src, dst, mu, sg, w = torch_ref.creature_tensors(cr, DEV)
    s0 = torch.as_tensor(cr["state"].astype(np.float64), device=DEV).requires_grad_(True)
    mu, sg, w = (t.clone().requires_grad_(True) for t in (mu, sg, w))
    xT = torch_ref.rollout(s0, kf, src, dst, mu, sg, w, steps)
    tgt = torch.as_tensor(np.asarray(target, np.float64).reshape(s0.shape), device=DEV)
    loss = ((xT - tgt) ** 2).mean()
    loss.backward()
    out = (float(loss), s0.grad.cpu().numpy(), mu.grad.cpu().numpy(), sg.grad.cpu().numpy(), w.grad.cpu().numpy())
    del xT, loss, s0
    torch.cuda.empty_cache()
    return out


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(np.max(np.abs(b)), 1e-300))


def test_config5_batched_search_against_torch_fp64(cn):
    T, ck = 32, 4
    crs, tgs = cn.creatures.search_batch_creatures(32, 512, 5, 0)
    tb = cn.CreatureBatch(tgs)
    for i, cr in enumerate(tgs):
        tb.load(i, cr)
    tb.run_steps(T)
    targets = [tb.creature_download(i) for i in range(len(tgs))]
    tb.close()
    b = cn.CreatureBatch(crs)
    for i, cr in enumerate(crs):
        b.load(i, cr)
        b.creature_set_target(i, targets[i])
    b.rollout_grad(T, ck)
    for i in (0, 17):
        loss_t, gs_t, gm_t, gsg_t, gw_t = _torch_grads(crs[i], targets[i], T)
        g = b.creature_grads(i)
        errs = {"loss": abs(b.creature_loss(i) - loss_t) / loss_t,
                "state0": _rel(g["state0"].reshape(gs_t.shape), gs_t), "mu": _rel(g["mu"], gm_t),
                "sigma": _rel(g["sigma"], gsg_t), "weights": _rel(g["weights"], gw_t)}
        print(f"512^2 C5K25 creature {i}, T={T}, ckpt {ck}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
        # bars about 10x the measured errors (loss <= 1e-7, gradients <= 1.6e-6)
        assert errs["loss"] <= 1e-6
        for k in ("state0", "mu", "sigma", "weights"):
            assert errs[k] <= 2e-5, (i, k, errs[k])
    # One device-side Adam iteration of the batch against single-creature engines.
    b.adam_step()
    b.rollout_grad(T, ck)
    flat = b.batch_growth()
    K = 25
    for i in (0, 17):
        e = cn.CreatureBatch([crs[i]])
        e.load(0, crs[i])
        e.creature_set_target(0, targets[i])
        e.rollout_grad(T, ck)
        e.adam_step()
        e.rollout_grad(T, ck)
        assert np.array_equal(e.batch_growth(), flat[3 * K * i:3 * K * (i + 1)]), i
        assert e.creature_loss(0) == b.creature_loss(i), i
        ge, gb = e.creature_grads(0), b.creature_grads(i)
        for k in ("state0", "mu", "sigma", "weights"):
            assert np.array_equal(ge[k], gb[k]), (i, k)
        e.close()
    b.close()
