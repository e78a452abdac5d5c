"""torch.autograd bridge for the checkpointed DiffusionLenia rollout (C ABI cn_rollout_forward /
cn_rollout_backward).  Sits beside the reference's PyTorch model (train/diff_lenia_org.py): where
that model is differentiated by autograd through torch.fft and F.unfold, this Function runs the
hand-written forward and adjoint HIP kernels and keeps only a checkpoint every `ckpt_every`
steps, recomputing each segment during the backward pass.

One engine holds one outstanding rollout: backward must come before the next forward on the
same engine (a backward whose forward was superseded raises instead of returning the newer
rollout's gradients).  Parameter gradients come back on the device and dtype of the parameters.
"""
import torch

from .engine import GpuMaceLenia


class DiffusionLeniaRollout(torch.autograd.Function):
    """x_T = rollout(x_0; mu, sigma, weights) with gradients for x_0, mu, sigma and weights."""

    @staticmethod
    def forward(ctx, state0, mu, sigma, weights, engine, steps, ckpt_every):
        if not isinstance(engine, GpuMaceLenia):
            raise TypeError("engine must be a GpuMaceLenia")
        if state0.dtype != torch.float32 or not state0.is_cuda:
            raise TypeError("state0 must be a float32 device tensor")
        C, n, K = engine._shapes[0]
        if state0.numel() != C * n * n:
            raise ValueError(f"state0 has {state0.numel()} elements, expected {C * n * n}")
        for name, t in (("mu", mu), ("sigma", sigma), ("weights", weights)):
            if t.numel() != K:
                raise ValueError(f"{name} has {t.numel()} elements, expected {K}")
        # From here on the engine is modified: an earlier forward's backward must not run on it.
        engine._rollout_id = getattr(engine, "_rollout_id", 0) + 1
        engine.set_growth(mu.detach().cpu().numpy(), sigma.detach().cpu().numpy(),
                          weights.detach().cpu().numpy())
        s0 = state0.detach().contiguous()
        torch.cuda.current_stream().synchronize()
        engine.creature_state_from_device(0, s0)
        engine.rollout_forward(steps, ckpt_every)
        out = torch.empty_like(s0)
        engine.creature_rollout_final(0, out)
        ctx.rollout_id = engine._rollout_id
        ctx.engine = engine
        ctx.K = mu.numel()
        ctx.param_meta = [(t.device, t.dtype, t.shape) for t in (mu, sigma, weights)]
        return out

    @staticmethod
    def backward(ctx, grad_out):
        eng = ctx.engine
        if getattr(eng, "_rollout_id", None) != ctx.rollout_id:
            raise RuntimeError("the engine ran another rollout forward since this one: call backward "
                               "before the next forward on the same engine")
        g = grad_out.detach().to(torch.float32).contiguous()
        torch.cuda.current_stream().synchronize()
        eng.rollout_backward([g])
        dev = g.device
        gs = torch.empty_like(g)
        gm, gsg, gw = (torch.empty(ctx.K, dtype=torch.float32, device=dev) for _ in range(3))
        eng.creature_grads(0, gs, gm, gsg, gw, device=True)
        gp = [t.to(device=d, dtype=dt).reshape(sh) for t, (d, dt, sh) in zip((gm, gsg, gw), ctx.param_meta)]
        return (gs, *gp, None, None, None)


def rollout(engine, state0, mu, sigma, weights, steps, ckpt_every=8):
    """Differentiable T-step rollout through the HIP engine (see DiffusionLeniaRollout)."""
    return DiffusionLeniaRollout.apply(state0, mu, sigma, weights, engine, int(steps), int(ckpt_every))
