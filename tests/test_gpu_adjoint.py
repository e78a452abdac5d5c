"""GPU checks of the hand-written adjoint kernels and the checkpointed rollout.

References: torch autograd of the float64 PyTorch restatement of the reference model
(oracle/torch_ref.py, restating train/diff_lenia_org.py:27-53), float64 central finite
differences of the same restatement (h = 1e-4 relative to the parameter scale), and a
dot-product (transpose) test."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import torch_ref  # noqa: E402

DEV = "cuda:0"


def _pair(cn, n, C, R, seed, patch):
    cr = cn.baseline_creature(n, num_channels=C, radius=R, seed=seed, patch=patch)
    tgt = cn.baseline_creature(n, num_channels=C, radius=R, seed=seed + 100, patch=patch)
    return cr, tgt


def _torch_grads(cr, target_state, steps, temp=1.0):
    kf, src, dst, mu, sg, w = torch_ref.creature_tensors(cr, DEV)
    s0 = torch.as_tensor(cr["state"].astype(np.float64), device=DEV).requires_grad_(True)
    mu, sg, w = (t.clone().requires_grad_(True) for t in (mu, sg, w))
    xT = torch_ref.rollout(s0, kf, src, dst, mu, sg, w, steps, temp)
    tgt = torch.as_tensor(target_state.astype(np.float64), device=DEV)
    loss = ((xT - tgt) ** 2).mean()
    loss.backward()
    return float(loss), s0.grad.cpu().numpy(), mu.grad.cpu().numpy(), sg.grad.cpu().numpy(), w.grad.cpu().numpy()


def _engine_grads(cn, cr, target_state, steps, ckpt):
    eng = cn.make_engine(cr)
    eng.creature_set_target(0, target_state)
    eng.rollout_grad(steps, ckpt)
    loss = eng.creature_loss(0)
    g = eng.creature_grads(0)
    final = eng.creature_rollout_final(0)
    eng.close()
    return loss, g, final


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("n,C,steps,ckpt", [(64, 3, 1, 1), (64, 3, 6, 2), (128, 2, 8, 3), (64, 4, 5, 5)])
def test_rollout_grad_matches_torch_fp64(cn, n, C, steps, ckpt):
    cr, tg = _pair(cn, n, C, R=max(4, n // 6), seed=n + C, patch=n // 2)
    tgt = tg["state"]
    loss_t, gs_t, gm_t, gsg_t, gw_t = _torch_grads(cr, tgt, steps)
    loss, g, _ = _engine_grads(cn, cr, tgt, steps, ckpt)
    errs = {"loss": abs(loss - loss_t) / loss_t,
            "state0": _rel(g["state0"].reshape(gs_t.shape), gs_t),
            "mu": _rel(g["mu"], gm_t), "sigma": _rel(g["sigma"], gsg_t), "weights": _rel(g["weights"], gw_t)}
    print(f"n={n} C={C} T={steps} ckpt={ckpt}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    # bars about 10x the measured errors (loss <= 1.7e-8, gradients <= 8.1e-7)
    assert errs["loss"] <= 2e-7
    for k in ("state0", "mu", "sigma", "weights"):
        assert errs[k] <= 1e-5, (k, errs[k])


def test_rollout_grad_with_shifted_kernels(cn):
    """Kernels whose spectra are not real and even (translated kernels) take the full complex
    path in the adjoint column pass (product with conj(K^)); checked against torch FP64."""
    from test_gpu_step import shift_kernels
    cr, tg = _pair(cn, 64, 3, R=10, seed=21, patch=32)
    shift_kernels(cr, [0, 4, 8], dx=2, dy=5)
    tgt = tg["state"]
    loss_t, gs_t, gm_t, gsg_t, gw_t = _torch_grads(cr, tgt, 4)
    loss, g, _ = _engine_grads(cn, cr, tgt, 4, 2)
    errs = {"loss": abs(loss - loss_t) / loss_t,
            "state0": _rel(g["state0"].reshape(gs_t.shape), gs_t),
            "mu": _rel(g["mu"], gm_t), "sigma": _rel(g["sigma"], gsg_t), "weights": _rel(g["weights"], gw_t)}
    print("shifted kernels: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    # measured: loss 2.5e-9, gradients <= 1.7e-6
    assert errs["loss"] <= 1e-7
    for k in ("state0", "mu", "sigma", "weights"):
        assert errs[k] <= 2e-5, (k, errs[k])


@pytest.mark.parametrize("maps", ["c1k1", "ragged_c3k5", "c4k2"])
def test_rollout_grad_ragged_maps(cn, maps):
    """Adjoint with a source channel that feeds no kernel, destinations that receive none and
    repeated (source, destination) pairs, against torch FP64 autograd."""
    from test_gpu_step import MAPS
    src, dst = MAPS[maps]
    C = max(max(src), max(dst)) + 1
    mk = lambda seed: cn.generate_creature(64, C, len(src), np.array(src), np.array(dst), 9, patch=32, seed=seed,
                                           creature=3)
    cr, tg = mk(31), mk(131)
    tgt = tg["state"]
    loss_t, gs_t, gm_t, gsg_t, gw_t = _torch_grads(cr, tgt, 4)
    loss, g, _ = _engine_grads(cn, cr, tgt, 4, 2)
    errs = {"loss": abs(loss - loss_t) / loss_t,
            "state0": _rel(g["state0"].reshape(gs_t.shape), gs_t),
            "mu": _rel(g["mu"], gm_t), "sigma": _rel(g["sigma"], gsg_t), "weights": _rel(g["weights"], gw_t)}
    print(f"{maps}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    # measured: loss <= 8e-9, gradients <= 3.5e-7
    assert errs["loss"] <= 1e-7
    for k in ("state0", "mu", "sigma", "weights"):
        assert errs[k] <= 5e-6, (k, errs[k])


def test_checkpoint_stride_does_not_change_gradients(cn):
    cr, tg = _pair(cn, 64, 3, R=10, seed=5, patch=32)
    ref = None
    for ck in (1, 2, 3, 7):
        _, g, _ = _engine_grads(cn, cr, tg["state"], 7, ck)
        if ref is None:
            ref = g
        else:
            for k in ("state0", "mu", "sigma", "weights"):
                assert np.array_equal(g[k], ref[k]), (ck, k)  # recompute is bit-identical


def test_rollout_final_equals_run_steps(cn):
    cr, tg = _pair(cn, 128, 3, R=12, seed=8, patch=64)
    _, _, final = _engine_grads(cn, cr, tg["state"], 9, 4)
    eng = cn.make_engine(cr)
    eng.run_steps(9)
    assert np.array_equal(final, eng.download_all_channels())
    eng.close()


def test_adjoint_dot_product_vs_fp64_finite_differences(cn):
    """<J v, w> (J v by float64 central differences of one reference step) = <v, J^T w> (HIP)."""
    cr, _ = _pair(cn, 64, 3, R=9, seed=21, patch=32)
    kf, src, dst, mu, sg, w_ = torch_ref.creature_tensors(cr, DEV)
    s0 = torch.as_tensor(cr["state"].astype(np.float64), device=DEV)
    rng = np.random.default_rng(0)
    v = torch.as_tensor(rng.standard_normal(s0.shape), device=DEV)
    wv = rng.standard_normal(s0.shape).astype(np.float32)
    h = 1e-4
    jv = (torch_ref.step(s0 + h * v, kf, src, dst, mu, sg, w_) - torch_ref.step(s0 - h * v, kf, src, dst, mu, sg, w_)) / (2 * h)
    lhs = float((jv.cpu().numpy() * wv).sum())
    eng = cn.make_engine(cr)
    eng.rollout_forward(1, 1)
    gw = torch.as_tensor(wv, device=DEV).contiguous()
    eng.rollout_backward([gw])
    jtw = eng.creature_grads(0)["state0"].reshape(s0.shape)
    rhs = float((v.cpu().numpy() * jtw).sum())
    eng.close()
    print(f"dot-product test: <Jv,w> = {lhs:.9e}, <v,J^T w> = {rhs:.9e}, rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_param_grads_vs_fp64_finite_differences(cn):
    cr, tg = _pair(cn, 64, 3, R=9, seed=33, patch=32)
    steps = 4
    kf, src, dst, mu, sg, w_ = torch_ref.creature_tensors(cr, DEV)
    s0 = torch.as_tensor(cr["state"].astype(np.float64), device=DEV)
    tgt = torch.as_tensor(tg["state"].astype(np.float64), device=DEV)

    def loss_of(m, s, w):
        return float(((torch_ref.rollout(s0, kf, src, dst, m, s, w, steps) - tgt) ** 2).mean())

    _, g, _ = _engine_grads(cn, cr, tg["state"], steps, 2)
    for name, idx, scale in (("mu", 0, 0.1), ("sigma", 4, 0.05), ("weights", 8, 0.3)):
        h = 1e-4 * scale
        args_p = [mu.clone(), sg.clone(), w_.clone()]
        args_m = [mu.clone(), sg.clone(), w_.clone()]
        j = ("mu", "sigma", "weights").index(name)
        args_p[j][idx] += h
        args_m[j][idx] -= h
        fd = (loss_of(*args_p) - loss_of(*args_m)) / (2 * h)
        print(f"d loss / d {name}[{idx}]: HIP {g[name][idx]:.6e}  FD(fp64) {fd:.6e}  rel {abs(g[name][idx] - fd) / abs(fd):.2e}")
        assert abs(g[name][idx] - fd) <= 1e-5 * abs(fd)  # agree to 7 significant digits


def test_split_backward_unaffected_by_conv_readback(cn):
    """cn_rollout_backward after cn_rollout_forward recomputes every segment's intermediates, so a
    call that reuses the intermediate buffers in between (the conv readback) changes nothing; and
    the split path gives the rollout graph's gradients (which reuse the forward's last segment)."""
    cr, tg = _pair(cn, 64, 3, R=9, seed=43, patch=32)
    target = tg["state"].reshape(-1).astype(np.float32)

    def split(readback):
        eng = cn.make_engine(cr)
        eng.run_steps(2)  # current state in slot 0, its predecessor kept in slot 1
        eng.rollout_forward(6, 4)
        final = eng.creature_rollout_final(0)
        if readback:
            eng.download_conv_buffer()
        gw = np.float32(2.0 / final.size) * (final - target)
        eng.rollout_backward([torch.as_tensor(gw, device=DEV).contiguous()])
        g = eng.creature_grads(0)
        eng.close()
        return g

    a, b = split(False), split(True)
    for k in ("state0", "mu", "sigma", "weights"):
        assert np.array_equal(a[k], b[k]), k
    eng = cn.make_engine(cr)
    eng.run_steps(2)
    eng.creature_set_target(0, tg["state"])
    eng.rollout_grad(6, 4)
    c = eng.creature_grads(0)
    eng.close()
    for k in ("state0", "mu", "sigma", "weights"):
        assert _rel(a[k], c[k].astype(np.float64)) <= 1e-6, k


def test_autograd_function_matches_internal_mse(cn):
    from cella_nova_amd.autograd import rollout
    cr, tg = _pair(cn, 64, 3, R=9, seed=41, patch=32)
    eng = cn.make_engine(cr)
    s0 = torch.as_tensor(cr["state"], device=DEV).requires_grad_(True)
    mu, sg, w = (torch.as_tensor(cr[k], device=DEV).requires_grad_(True) for k in ("mu", "sigma", "weights"))
    xT = rollout(eng, s0, mu, sg, w, steps=5, ckpt_every=2)
    tgt = torch.as_tensor(tg["state"], device=DEV)
    loss = ((xT - tgt) ** 2).mean()
    loss.backward()
    _, g, final = _engine_grads(cn, cr, tg["state"], 5, 2)
    assert np.array_equal(xT.detach().cpu().numpy().reshape(-1), final)
    assert _rel(s0.grad.cpu().numpy().reshape(-1), g["state0"].astype(np.float64)) <= 1e-5
    for name, t in (("mu", mu), ("sigma", sg), ("weights", w)):
        assert _rel(t.grad.cpu().numpy(), g[name].astype(np.float64)) <= 1e-5
    # and directly against torch autograd of the float64 restatement
    loss_t, gs_t, gm_t, gsg_t, gw_t = _torch_grads(cr, tg["state"], 5)
    errs = {"loss": abs(float(loss) - loss_t) / loss_t, "state0": _rel(s0.grad.cpu().numpy(), gs_t),
            "mu": _rel(mu.grad.cpu().numpy(), gm_t), "sigma": _rel(sg.grad.cpu().numpy(), gsg_t),
            "weights": _rel(w.grad.cpu().numpy(), gw_t)}
    print("autograd.rollout vs torch FP64 autograd: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    # measured: loss 1.6e-8, gradients <= 2.2e-7
    assert errs["loss"] <= 2e-7
    for k in ("state0", "mu", "sigma", "weights"):
        assert errs[k] <= 5e-6, (k, errs[k])
    eng.close()


def test_differentiable_rollout_1024_c3k9_t64(cn):
    """Differentiable rollout at 1024^2, C3K9, T=64, checkpoint every 8: MSE to a second seeded
    creature's state after 64 steps; gradients for the 3x128x128 seed patch and mu/sigma/w
    against float64 torch autograd of the reference restatement, and against FP64 central finite
    differences (h = 1e-4) of the same restatement for all 27 growth parameters and 8 seed pixels."""
    n, T = 1024, 64
    cr, tg = _pair(cn, n, 3, R=13, seed=0, patch=128)
    eng = cn.make_engine(tg)
    eng.run_steps(T)
    target = eng.download_all_channels().reshape(3, n, n)
    eng.close()
    loss_t, gs_t, gm_t, gsg_t, gw_t = _torch_grads(cr, target, T)
    loss, g, _ = _engine_grads(cn, cr, target, T, 8)
    sl = slice(n // 2 - 64, n // 2 + 64)
    patch_t = gs_t[:, sl, sl]
    patch = g["state0"].reshape(3, n, n)[:, sl, sl]
    errs = {"loss": abs(loss - loss_t) / loss_t, "seed_patch": _rel(patch, patch_t), "mu": _rel(g["mu"], gm_t),
            "sigma": _rel(g["sigma"], gsg_t), "weights": _rel(g["weights"], gw_t)}
    print("1024^2 C3K9 T=64 ckpt 8, engine vs torch FP64 autograd: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))

    # FP64 central finite differences (h = 1e-4) of the float64 restatement at the same shape:
    # every mu, sigma and weight, and the 8 seed-patch pixels with the largest gradient.
    kf, src, dst, mu, sg, w = torch_ref.creature_tensors(cr, DEV)
    s0 = torch.as_tensor(cr["state"].astype(np.float64), device=DEV)
    tgt = torch.as_tensor(target.astype(np.float64), device=DEV)

    def loss_of(s, m, sgm, ww):
        with torch.no_grad():
            return float(((torch_ref.rollout(s, kf, src, dst, m, sgm, ww, T) - tgt) ** 2).mean())

    h = 1e-4
    fd = {}
    for j, name in enumerate(("mu", "sigma", "weights")):
        vals = []
        for k in range(9):
            args_p, args_m = [mu.clone(), sg.clone(), w.clone()], [mu.clone(), sg.clone(), w.clone()]
            args_p[j][k] += h
            args_m[j][k] -= h
            vals.append((loss_of(s0, *args_p) - loss_of(s0, *args_m)) / (2 * h))
        fd[name] = np.array(vals)
    flat = np.argsort(-np.abs(patch.reshape(-1)))[:8]
    pix = [np.unravel_index(f, patch.shape) for f in flat]
    vals = []
    for (c, y, x) in pix:
        sp, sm = s0.clone(), s0.clone()
        sp[c, n // 2 - 64 + y, n // 2 - 64 + x] += h
        sm[c, n // 2 - 64 + y, n // 2 - 64 + x] -= h
        vals.append((loss_of(sp, mu, sg, w) - loss_of(sm, mu, sg, w)) / (2 * h))
    fd["pixels"] = np.array(vals)
    eng_v = {"mu": g["mu"], "sigma": g["sigma"], "weights": g["weights"], "pixels": patch.reshape(-1)[flat]}
    tor_v = {"mu": gm_t, "sigma": gsg_t, "weights": gw_t, "pixels": patch_t.reshape(-1)[flat]}
    fd_e = {k: _rel(eng_v[k], fd[k]) for k in fd}
    fd_t = {k: _rel(tor_v[k], fd[k]) for k in fd}
    print("engine vs FP64 central differences: " + ", ".join(f"{k} {v:.2e}" for k, v in fd_e.items()))
    print("torch FP64 autograd vs FP64 central differences: " + ", ".join(f"{k} {v:.2e}" for k, v in fd_t.items()))
    # measured: engine vs autograd loss 3.8e-8, gradients <= 2.2e-6; engine vs differences
    # <= 3.7e-6; autograd vs differences 3.9e-6 for mu and sigma (the differences' own
    # truncation error), <= 2.5e-8 for the weights and pixels
    assert errs["loss"] <= 1e-6
    for k in ("seed_patch", "mu", "sigma", "weights"):
        assert errs[k] <= 1e-5, (k, errs[k])
    for k in fd:
        assert fd_t[k] <= 2e-5, ("torch", k, fd_t[k])
        assert fd_e[k] <= 2e-5, ("engine", k, fd_e[k])


def test_rollout_grad_maximum_channels_and_kernels(cn):
    """C = 16, K = 256 (all pairs): rollout gradients against torch FP64 autograd."""
    cr, tg = _pair(cn, 32, 16, R=5, seed=41, patch=16)
    loss_t, gs_t, gm_t, gsg_t, gw_t = _torch_grads(cr, tg["state"], 3)
    loss, g, _ = _engine_grads(cn, cr, tg["state"], 3, 2)
    errs = {"loss": abs(loss - loss_t) / loss_t,
            "state0": _rel(g["state0"].reshape(gs_t.shape), gs_t),
            "mu": _rel(g["mu"], gm_t), "sigma": _rel(g["sigma"], gsg_t), "weights": _rel(g["weights"], gw_t)}
    print("32^2 C16K256 T=3 ckpt=2: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    # measured: loss 1.2e-8, gradients <= 3e-7
    assert errs["loss"] <= 2e-7
    for k in ("state0", "mu", "sigma", "weights"):
        assert errs[k] <= 5e-6, (k, errs[k])


def test_autograd_function_cpu_parameters_and_superseded_forward(cn):
    """Growth parameters given as CPU float64 tensors get their gradients as CPU float64 tensors;
    a backward whose forward was superseded by another forward on the same engine raises."""
    from cella_nova_amd.autograd import rollout
    cr, tg = _pair(cn, 64, 3, R=9, seed=43, patch=32)
    eng = cn.make_engine(cr)
    s0 = torch.as_tensor(cr["state"], device=DEV).requires_grad_(True)
    mu, sg, w = (torch.as_tensor(cr[k].astype(np.float64)).requires_grad_(True) for k in ("mu", "sigma", "weights"))
    tgt = torch.as_tensor(tg["state"], device=DEV)
    loss = ((rollout(eng, s0, mu, sg, w, steps=3, ckpt_every=2) - tgt) ** 2).mean()
    loss.backward()
    _, g, _ = _engine_grads(cn, cr, tg["state"], 3, 2)
    for name, t in (("mu", mu), ("sigma", sg), ("weights", w)):
        assert t.grad.device.type == "cpu" and t.grad.dtype == torch.float64
        assert _rel(t.grad.numpy(), g[name].astype(np.float64)) <= 1e-5
    first = ((rollout(eng, s0, mu, sg, w, steps=3, ckpt_every=2) - tgt) ** 2).mean()
    second = ((rollout(eng, s0, mu, sg, w, steps=4, ckpt_every=2) - tgt) ** 2).mean()
    with pytest.raises(RuntimeError, match="another rollout forward"):
        first.backward()
    second.backward()
    eng.close()


@pytest.mark.parametrize("n", [256, 1024])
def test_rollout_grad_mixed_row_geometries(cn, n):
    """Forward and inverse row transforms with different values per thread (different row-block
    counts): the parameter-gradient reduction reads exactly the partials the growth adjoint of
    this plan wrote, also after a plan with more row blocks left its partials in the buffer."""
    cr, tg = _pair(cn, n, 3, R=max(4, n // 20), seed=47, patch=min(128, n // 2))
    loss_t, gs_t, gm_t, gsg_t, gw_t = _torch_grads(cr, tg["state"], 3)
    eng = cn.make_engine(cr)
    eng.creature_set_target(0, tg["state"])
    for fe, ie in ((8, 8), (16, 8), (8, 16), (16, 16), (8, 16)):
        eng.set_plan(n, fwd_e=fe, inv_e=ie)
        eng.rollout_grad(3, 2)
        g = eng.creature_grads(0)
        errs = {"state0": _rel(g["state0"].reshape(gs_t.shape), gs_t), "mu": _rel(g["mu"], gm_t),
                "sigma": _rel(g["sigma"], gsg_t), "weights": _rel(g["weights"], gw_t)}
        print(f"n={n} fwd_e={fe} inv_e={ie}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
        # measured: <= 4.5e-7 at 256^2; at 1024^2 the FP32 parameter sums over 2^20 sites per
        # step reach 9.5e-6
        for k, v in errs.items():
            assert v <= (5e-6 if n <= 256 else 1e-4), (fe, ie, k, v)
    eng.close()


# (channels, source map, destination map): channel 1 feeds no kernel; channel 0 feeds none and
# channel 1 receives none; then random maps
SEG_GRAD_CASES = [(3, [0, 0, 2, 2], [1, 2, 0, 1]), (2, [1], [0]), None, None]


@pytest.mark.parametrize("case", range(len(SEG_GRAD_CASES)))
def test_segmented_adjoint_random_creatures(cn, case):
    """512^2 creatures on the segmented column pass (per-kernel radii 3-30, ragged maps with
    sources that feed no kernel, whose adjoint segment spectra are zero): rollout gradients
    against torch FP64 autograd and against the full-length adjoint of the same creature."""
    n, steps = 512, 3
    rng = np.random.default_rng(500 + case)
    if SEG_GRAD_CASES[case] is None:
        C = int(rng.integers(1, 5))
        K = int(rng.integers(1, 2 * C + 2))
        src, dst = rng.integers(0, C, K), rng.integers(0, C, K)
    else:
        C, src, dst = SEG_GRAD_CASES[case]
        src, dst, K = np.array(src), np.array(dst), len(src)
    radii = rng.integers(3, 31, K).astype(np.uint32)
    mk = lambda seed: cn.generate_creature(n, C, K, src, dst, radii, patch=128, seed=seed, creature=4)
    cr, tg = mk(60 + case), mk(160 + case)
    loss_t, gs_t, gm_t, gsg_t, gw_t = _torch_grads(cr, tg["state"], steps)
    got = {}
    for full in (0, 1):
        eng = cn.make_engine(cr)
        segs = eng.info()["segments"]
        assert segs > 0
        eng.set_plan(n, col_full=full)
        eng.creature_set_target(0, tg["state"])
        eng.rollout_grad(steps, 2)
        got[full] = (eng.creature_loss(0), eng.creature_grads(0))
        eng.close()
    loss, g = got[0]
    errs = {"loss": abs(loss - loss_t) / loss_t,
            "state0": _rel(g["state0"].reshape(gs_t.shape), gs_t),
            "mu": _rel(g["mu"], gm_t), "sigma": _rel(g["sigma"], gsg_t), "weights": _rel(g["weights"], gw_t)}
    gf = got[1][1]
    vs_full = {k: _rel(g[k], gf[k].astype(np.float64)) for k in ("state0", "mu", "sigma", "weights")}
    print(f"case {case}: C{C} K{K} src {src.tolist()} dst {dst.tolist()} radii {radii.tolist()} segments {segs}: "
          + ", ".join(f"{k} {v:.2e}" for k, v in errs.items())
          + " | vs full-length: " + ", ".join(f"{k} {v:.2e}" for k, v in vs_full.items()))
    # measured: loss <= 5.1e-9, gradients <= 1.3e-6 against torch and <= 2.4e-6 against the
    # full-length adjoint
    assert errs["loss"] <= 5e-8
    for k in ("state0", "mu", "sigma", "weights"):
        assert errs[k] <= 1.5e-5, (k, errs[k])
        assert vs_full[k] <= 3e-5, ("full-length", k, vs_full[k])
