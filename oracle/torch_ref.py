"""ORACLE - test infrastructure only (imported by tests/ as the checker; the product never
imports it).

PyTorch restatement of the reference model's step, DiffusionLenia.step / compute_affinity
(train/diff_lenia_org.py:27-53) with MCLenia.get_fftconv and growth (train/lenia_org.py:232-275),
per kernel instead of the reference's dense (B, C, C, H, W) broadcast: u_k = Re(ifft2(fft2(s_src)
K^_k)), G = 2 exp(-((u - mu)^2 / sigma^2) / 2) - 1, aff_c = sum_{k: dst = c} w_k G_k,
A = exp(temp aff), Z = circular 3x3 sum of A (F.pad circular + unfold), s' = A * box3(s / Z).
Differentiable by torch autograd in any dtype, on any device: the gradient reference for the
HIP adjoint kernels (run in float64).
"""
import torch


def box3(x):
    s = torch.zeros_like(x)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            s = s + torch.roll(x, shifts=(dy, dx), dims=(-2, -1))
    return s


def step(state, kfft, src, dst, mu, sigma, w, temp=1.0):
    """state (C, H, W) real; kfft (K, H, W) complex; src/dst (K,) long; mu/sigma/w (K,)."""
    S = torch.fft.fft2(state)
    U = torch.fft.ifft2(S[src] * kfft).real
    G = 2 * torch.exp(-((U - mu[:, None, None]) ** 2 / sigma[:, None, None] ** 2) / 2) - 1
    aff = torch.zeros_like(state).index_add(0, dst, G * w[:, None, None])
    A = torch.exp(temp * aff)
    Z = box3(A)
    return A * box3(state / Z)


def rollout(state0, kfft, src, dst, mu, sigma, w, steps, temp=1.0):
    s = state0
    for _ in range(steps):
        s = step(s, kfft, src, dst, mu, sigma, w, temp)
    return s


def creature_tensors(cr, device, dtype=torch.float64):
    """Tensors of a generated creature (creatures.generate_creature output) for this module."""
    n, K = cr["n"], cr["num_kernels"]
    cdt = torch.complex128 if dtype == torch.float64 else torch.complex64
    kf = torch.as_tensor(cr["kernels_fft"].reshape(K, n, n)).to(device=device, dtype=cdt)
    src = torch.as_tensor(cr["src"].astype("int64"), device=device)
    dst = torch.as_tensor(cr["dst"].astype("int64"), device=device)
    p = {k: torch.as_tensor(cr[k].astype("float64")).to(device=device, dtype=dtype) for k in ("mu", "sigma", "weights")}
    return kf, src, dst, p["mu"], p["sigma"], p["weights"]
