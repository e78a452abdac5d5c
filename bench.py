"""Benchmark of the DiffusionLenia step engine on one MI355X.

Workload: 2048 x 2048 periodic grid, 3 channels, 9 kernels (src c0=[0,1,2]*3, dst
c1=[0,0,0,1,1,1,2,2,2]), Lenia ring kernels of radius 26, fp32, temp 1, synthetic seeded creature.
A step is one full convolution -> growth -> diffusion -> copy pass over the grid; inputs are
resident in HBM before the timed region.  Prints one JSON line (rank 0):
  value      cell-steps/s (grid sites advanced per second, all channels)
  roofline   dominant kernel: algorithmic bytes per launch / average launch time (HIP events
             on the engine's stream), against 8 TB/s HBM
  cpu_baseline  the C++ restatement of the reference's CPU implementation (ml-rs-example),
             single thread, on a bounded sample of the 1024^2 C3K9 R=13 creature.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

METRIC = "cell-steps/s 2048² fp32; batched fwd+bwd grad-steps/s vs plain at 1e-5 rel err"
HBM_PEAK_GBS = 8000.0
KERNELS = ["row_fwd_r2c", "col_fused", "row_inv_growth", "diffusion"]


def algorithmic_bytes_per_cell(C, K, compact=0, seg_table=None):
    """Bytes each kernel must move per grid site (all channels), by its inputs and outputs:
    row_fwd reads C f32 planes, writes C half spectra (N/2 complex per row = 4 B/site);
    col_fused reads C half spectra + K kernel spectra, writes K half spectra; a kernel spectrum
    is a complex half spectrum (4 B/site) or, for the `compact` kernels whose spectrum is real
    and even in ky, a real quarter (1 B/site); with the segmented column pass it is the kernel's
    segment table instead (`seg_table` B/site per kernel: (M/2+1)(N/2+1) f32 over N^2 sites);
    row_inv_growth reads K half spectra, writes C affinity planes;
    diffusion reads C affinity + C state planes, writes C state planes."""
    kern = K * seg_table if seg_table is not None else 4 * (K - compact) + compact
    return {"row_fwd_r2c": 8 * C, "col_fused": 4 * C + 4 * K + kern,
            "row_inv_growth": 4 * K + 4 * C, "diffusion": 12 * C}


def seg_table_bytes_per_cell(info):
    """Segment-table bytes per grid site per kernel, or None without the segmented pass."""
    m, n = info.get("segment_length", 0), info["n"]
    return (m // 2 + 1) * (n // 2 + 1) * 4 / (n * n) if info.get("segments", 0) and m else None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(sample_steps):
    """ml-rs-example restatement (oracle/cpu_ref.cpp, f32, one thread) on 1024^2 C3K9 R=13,
    pinned to one core of the host (the reference's CPU example is single-threaded)."""
    orc = ge.load_oracle()
    pkg = ge.load_package()
    cr = pkg.baseline_creature(1024, num_channels=3, radius=13, seed=0)
    allowed = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else None
    core = min(allowed) if allowed else None
    if core is not None:
        os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        orc.run_f32(cr["state"], cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"],
                    steps=sample_steps)
        dt = time.perf_counter() - t0
    finally:
        if core is not None:
            os.sched_setaffinity(0, allowed)
    pin = f"pinned to core {core} of {len(allowed)}" if core is not None else "unpinned"
    return {"value": 1024 * 1024 * sample_steps / dt, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"{sample_steps} steps of 1024x1024 C3K9 R=13 (ml-rs-example restatement, f32, "
                      f"1 thread {pin}, {dt:.1f} s; host {_cpu_model()}, nproc {os.cpu_count()})"}


def _adam(p, g, m, v, t, lr=5e-3, b1=0.9, b2=0.999, eps=1e-8):
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    return p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)


def _sanitize(mu, sigma, w):
    """Parameter box of the reference's LeniaParams._sanitize (train/utils/leniaparams.py:351-360)."""
    return np.clip(mu, 0.0, 2.0), np.maximum(sigma, 1e-3), np.maximum(w, 0.0)


def torch_reference_rate(crs, targets, T, iters, device="cuda:0"):
    """The reference's own GPU gradient path, one creature at a time: the PyTorch DiffusionLenia
    step (train/diff_lenia_org.py:27-53 with MCLenia.get_fftconv / growth, lenia_org.py:232-275:
    dense (C, C) kernel broadcast, F.pad circular + F.unfold 3x3 sums) in fp32, autograd through
    T steps (no checkpointing, as in the reference), MSE to the target, torch.optim.Adam on
    mu/sigma/weights.  Returns (grad-steps/s, mean final loss)."""
    import torch
    import torch.nn.functional as F

    def step(state, kd, mu, sigma, w):
        C, H, W = state.shape
        S = torch.fft.fft2(state)[:, None]                       # (C_src, 1, H, W)
        U = torch.fft.ifft2(S * kd).real                          # (C_src, C_dst, H, W)
        G = 2 * torch.exp(-((U - mu[..., None, None]) ** 2 / sigma[..., None, None] ** 2) / 2) - 1
        A = torch.exp((G * w[..., None, None]).sum(dim=0))[None]  # (1, C, H, W), temp = 1
        Z = F.unfold(F.pad(A, (1, 1, 1, 1), mode="circular"), kernel_size=(3, 3)).reshape(1, C, 9, H, W).sum(2)
        q = F.pad(state[None] / Z, (1, 1, 1, 1), mode="circular")
        q = F.unfold(q, kernel_size=(3, 3)).reshape(1, C, 9, H, W)
        return (A[:, :, None] * q).sum(dim=2)[0]

    def dense(cr, key):
        C = cr["num_channels"]
        a = np.zeros((C, C), np.float32)
        a[cr["src"], cr["dst"]] = cr[key]
        return torch.tensor(a, device=device)

    setups = []
    for cr, tg in zip(crs, targets):
        C, n = cr["num_channels"], cr["n"]
        kd = np.zeros((C, C, n, n), np.complex64)
        kd[cr["src"], cr["dst"]] = np.asarray(cr["kernels_fft"]).reshape(-1, n, n)
        params = [dense(cr, k).requires_grad_(True) for k in ("mu", "sigma", "weights")]
        setups.append((torch.tensor(cr["state"], device=device), torch.tensor(kd, device=device),
                       torch.tensor(np.asarray(tg).reshape(C, n, n), device=device), params,
                       torch.optim.Adam(params, lr=5e-3)))
    losses = []
    t0 = None
    for it in range(iters + 1):  # iteration 0 warms up cuFFT plans and the allocator (untimed)
        if it == 1:
            if device.startswith("cuda"):
                torch.cuda.synchronize()
            t0 = time.perf_counter()
        for s0, kd, tgt, params, opt in setups:
            opt.zero_grad()
            s = s0
            for _ in range(T):
                s = step(s, kd, *params)
            loss = ((s - tgt) ** 2).mean()
            loss.backward()
            opt.step()
            with torch.no_grad():
                params[0].clamp_(0.0, 2.0)
                params[1].clamp_(min=1e-3)
                params[2].clamp_(min=0.0)
            if it == iters:
                losses.append(float(loss.detach()))
    if device.startswith("cuda"):
        torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return len(setups) * T * iters / dt, float(np.mean(losses))


def grad_step_bytes_per_cell(C, K, T, ckpt):
    """Algorithmic bytes per grad-step (one creature advanced and back-propagated through one
    step) per grid site, for a T-step rollout with a checkpoint every `ckpt` steps, all kernel
    spectra compact: the inputs and outputs of every launch of the rollout, divided by T.
      forward step   28C + 9K (as the step), + 4K for each step of the last segment (writes u)
      recompute step 28C + 13K (the other segments' steps again, writing u), minus the
                     diffusion (12C) of each such segment's last step
      adjoint step   diffusion adjoint 20C (A, state, g in; dsd, gaff out), growth adjoint with
                     the row transform 12K (u, gaff in; spectrum out), column adjoint 4C + 5K,
                     row inverse 12C (spectrum, dsd in; gradient out)
      loss           12C once (final state and target in, gradient out)"""
    ck = min(ckpt, T)
    nck = (T + ck - 1) // ck
    last = T - (nck - 1) * ck
    fwd = T * (28 * C + 9 * K) + last * 4 * K
    rec = (T - last) * (28 * C + 13 * K) - (nck - 1) * 12 * C
    adj = T * (36 * C + 17 * K)
    return (fwd + rec + adj + 12 * C) / T


def _apply_plans(eng, plans):
    """Launch plans (cn_plan dicts, e.g. the "plans" a tuned run printed) for the engine's size
    classes, instead of tuning: keeps the tuner's candidate launches out of a profile."""
    by_n = {int(p["n"]): p for p in plans}
    for cur in eng.plans():
        p = by_n.get(int(cur["n"]))
        if p is not None:
            eng.set_plan(**{k: int(v) for k, v in p.items()})
    return eng.plans()


def batched_grad(pkg, count=32, n=512, C=5, T=32, ckpt=4, iters=50, plain_iters=None, seed=0, torch_iters=2,
                 plans=None):
    """Batched gradient-based creature search: `count` creatures (C channels, K = C^2 kernels,
    radii U{10..18}), T-step rollouts with a checkpoint every `ckpt` steps, MSE to the state a
    second seeded creature reaches after T steps, Adam on mu/sigma/weights with the sanitize box
    of LeniaParams._sanitize.  Rate in grad-steps/s (one creature advanced and back-propagated
    through one step):
      batched  one engine for all creatures: per iteration one rollout graph and one device-side
               Adam launch, no host round trip;
      plain    one engine per creature (the same kernels, graphs and device-side Adam, one tuned
               plan), iterated creature by creature;
      torch    the reference's PyTorch path per creature (torch_reference_rate)."""
    crs, tgs = pkg.creatures.search_batch_creatures(count, n, C, seed)
    K = C * C
    tb = pkg.CreatureBatch(tgs)
    for i, cr in enumerate(tgs):
        tb.load(i, cr)
    tb.run_steps(T)
    targets = [tb.creature_download(i) for i in range(count)]
    tb.close()

    def build(members, plan=None):
        eng = pkg.CreatureBatch([crs[i] for i in members])
        for j, i in enumerate(members):
            eng.load(j, crs[i])
            eng.creature_set_target(j, targets[i])
        if plan is None:
            plan = _apply_plans(eng, plans)[0] if plans else eng.autotune(3)[0]
        else:
            eng.set_plan(**plan)
        return eng, plan

    def run(engines, n_iter):
        """engines: (engine, member creatures); returns (seconds, final losses)."""
        for eng, members in engines:  # iteration 0 captures the graphs, allocates the Adam state
            eng.rollout_grad(T, ckpt)
            eng.adam_step()
            eng.adam_reset()         # ... and the timed iterations start from the initial state
            eng.batch_set_growth(np.concatenate([np.concatenate([crs[i][k] for k in ("mu", "sigma", "weights")])
                                                 for i in members]))
            eng.sync()
        t0 = time.perf_counter()
        for _ in range(n_iter):
            for eng, _ in engines:
                eng.rollout_grad(T, ckpt)
                eng.adam_step()
        for eng, _ in engines:
            eng.sync()
        dt = time.perf_counter() - t0
        return dt, [eng.creature_loss(j) for eng, members in engines for j in range(len(members))]

    batch, plan = build(list(range(count)))
    dt_b, loss_b = run([(batch, list(range(count)))], iters)
    batch.close()
    pi = iters if plain_iters is None else plain_iters
    rate_p, loss_p = None, [float("nan")]
    if pi:
        plain = [(build([i], plan)[0], [i]) for i in range(count)]
        dt_p, loss_p = run(plain, pi)
        for e, _ in plain:
            e.close()
        rate_p = count * T * pi / dt_p
    rate_b = count * T * iters / dt_b
    bpc = grad_step_bytes_per_cell(C, K, T, ckpt)
    gbps = bpc * n * n * rate_b / 1e9
    out = {"config": f"{count} creatures {n}x{n} C{C}K{K} R~U{{10..18}}, T={T}, ckpt {ckpt}, Adam on mu/sigma/w "
                     f"(device-side), {iters} iterations (plain loop: {pi}, PyTorch reference path: {torch_iters})",
           "grad_steps_per_s_batched": round(rate_b, 1),
           "grad_steps_per_s_plain": round(rate_p, 1) if rate_p else None,
           "ratio": round(rate_b / rate_p, 2) if rate_p else None, "host_round_trips_per_iteration": 0,
           "final_mse_mean_batched": float(np.mean(loss_b)), "final_mse_mean_plain": float(np.mean(loss_p)),
           "alg_bytes_per_grad_step": round(bpc * n * n), "alg_GBps": round(gbps, 1),
           "hbm_frac": round(gbps / HBM_PEAK_GBS, 4), "plan": plan}
    if torch_iters:
        rate_t, loss_t = torch_reference_rate(crs, targets, T, torch_iters)
        out.update({"grad_steps_per_s_torch_reference": round(rate_t, 1), "ratio_vs_torch_reference":
                    round(rate_b / rate_t, 2), "final_mse_mean_torch_reference": loss_t})
    return out


def batched_forward(pkg, steps=200, seed=0, plans=None):
    """256 creatures of uneven sizes (85 at 256^2, 85 at 512^2, 86 at 1024^2), C ~ U{2,3,4},
    K = C^2, radii U{8..20}, one engine and one graph per step for all of them.  The creatures
    are generated on the device (kernel spectra, growth parameters, seed); the host generator's
    time for the same creatures is sampled on a few of them for comparison."""
    specs = pkg.creatures.uneven_batch_specs(seed)
    sizes = [sp["n"] for sp in specs]
    t0 = time.perf_counter()
    b = pkg.CreatureBatch.generated(specs)
    b.sync()
    setup_dev = time.perf_counter() - t0
    # Host FP64 generator on the first 4 creatures of each size, scaled to the whole batch.
    t_host = 0.0
    for first in (0, 85, 170):
        t1 = time.perf_counter()
        for sp in specs[first:first + 4]:
            pkg.generate_creature(sp["n"], sp["num_channels"], sp["num_kernels"], sp["src"], sp["dst"],
                                  sp["radius"], patch=sp["patch"], seed=sp["seed"], creature=sp["creature"])
        t_host += (time.perf_counter() - t1) * (85 if first < 170 else 86) / 4
    plans = _apply_plans(b, plans) if plans else b.autotune(3)
    b.run_steps(5)
    b.sync()
    t0 = time.perf_counter()
    b.run_steps(steps)
    b.sync()
    dt = time.perf_counter() - t0
    b.close()
    cells = sum(n * n for n in sizes)
    # Algorithmic bytes per cell-step: 28 C + 9 K with every kernel on the compact real path.
    alg = sum((28 * sp["num_channels"] + 9 * sp["num_kernels"]) * sp["n"] ** 2 for sp in specs)
    return {"config": f"{len(specs)} creatures (85x256^2, 85x512^2, 86x1024^2), C~U{{2,3,4}}, K=C^2, R~U{{8..20}}, "
                      f"{steps} steps", "cell_steps_per_s": round(cells * steps / dt, 1),
            "ms_per_step": round(dt * 1e3 / steps, 4), "alg_GBps": round(alg * steps / dt / 1e9, 1),
            "hbm_frac": round(alg * steps / dt / 1e9 / HBM_PEAK_GBS, 4), "plans": plans,
            "setup_s_device_generator": round(setup_dev, 3), "setup_s_host_generator_est": round(t_host, 1)}


def rollout_grad(pkg, n=1024, R=13, T=64, ckpt=8, iters=5, torch_iters=1, seed=0, plans=None):
    """Differentiable rollout (config #4): one 1024^2 C3K9 creature, T steps with a checkpoint
    every `ckpt`, MSE to the state a second seeded creature reaches after T steps, gradients for
    the state and mu/sigma/w through the hand-written adjoint (one HIP graph per rollout).
    Reported as grad-steps/s and ms per rollout, next to the reference's PyTorch autograd path
    on the same creature (torch_reference_rate)."""
    cr = pkg.baseline_creature(n, num_channels=3, radius=R, seed=seed)
    tg = pkg.baseline_creature(n, num_channels=3, radius=R, seed=seed + 1)
    te = pkg.make_engine(tg)
    te.run_steps(T)
    target = te.download_all_channels().reshape(3, n, n)
    te.close()
    eng = pkg.make_engine(cr)
    plan = _apply_plans(eng, plans)[0] if plans else eng.autotune(5)[0]
    eng.creature_set_target(0, target)
    eng.rollout_grad(T, ckpt)  # captures the graph
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(iters):
        eng.rollout_grad(T, ckpt)
    eng.sync()
    dt = (time.perf_counter() - t0) / iters
    loss = eng.creature_loss(0)
    eng.close()
    gbps = grad_step_bytes_per_cell(3, 9, T, ckpt) * n * n * T / dt / 1e9
    out = {"config": f"{n}x{n} C3K9 R={R}, T={T}, checkpoint every {ckpt}, {iters} rollouts",
           "ms_per_rollout": round(dt * 1e3, 3), "grad_steps_per_s": round(T / dt, 1), "loss": loss,
           "alg_bytes_per_grad_step": round(grad_step_bytes_per_cell(3, 9, T, ckpt) * n * n),
           "alg_GBps": round(gbps, 1), "hbm_frac": round(gbps / HBM_PEAK_GBS, 4), "plan": plan}
    if torch_iters:
        rate_t, loss_t = torch_reference_rate([cr], [target], T, torch_iters)
        out.update({"grad_steps_per_s_torch_reference": round(rate_t, 1),
                    "ratio_vs_torch_reference": round(T / dt / rate_t, 2)})
    return out


def load_traffic(workload_tag, kernel):
    """Measured HBM bytes per launch from committed rocprofv3 PMC summaries, if present."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload_tag}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--radius", type=int, default=26)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-grad", action="store_true",
                    help="step2048: leave out the batched gradient-search measurement")
    ap.add_argument("--workload", default="step2048",
                    choices=["step2048", "batched-forward", "batched-grad", "rollout"],
                    help="step2048: the headline 2048^2 C3K9 step; the batched workloads print their own line")
    ap.add_argument("--iters", type=int, default=50, help="batched-grad: Adam iterations")
    ap.add_argument("--plain-iters", type=int, default=None,
                    help="batched-grad: iterations of the plain per-creature loop (default: --iters; 0: skip)")
    ap.add_argument("--torch-iters", type=int, default=2,
                    help="batched-grad: iterations of the PyTorch reference path (0: skip)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--plan", default="",
                    help="fixed launch plan fwd_e,inv_e,col_e,col_tc,diff_th[,col_split[,col_full]] (no autotune)")
    ap.add_argument("--plans", default="",
                    help="batched workloads: JSON list of launch plans (as printed under \"plans\") instead "
                         "of tuning")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run a device-to-device copy of this many MiB (PMC calibration only)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this benchmark runs on one GPU only (--gpus 1); creature batches shard across "
              "GPUs without communication, one process and one CreatureBatch per GPU", file=sys.stderr)
        sys.exit(2)

    pkg = ge.load_package()
    plans = json.loads(args.plans) if args.plans else None
    if args.copy_calibration_mb:
        import torch
        a = torch.ones(args.copy_calibration_mb << 18, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "batched-grad":
        print(json.dumps({"workload": "batched-grad", **batched_grad(pkg, iters=args.iters, torch_iters=args.torch_iters,
                                                                      plain_iters=args.plain_iters,
                                                                      plans=plans)}), flush=True)
        return
    if args.workload == "rollout":
        print(json.dumps({"workload": "rollout", **rollout_grad(pkg, torch_iters=args.torch_iters, plans=plans)}),
              flush=True)
        return
    if args.workload == "batched-forward":
        print(json.dumps({"workload": "batched-forward", **batched_forward(pkg, steps=args.steps, plans=plans)}),
              flush=True)
        return
    n, C, R = args.n, 3, args.radius
    cr = pkg.baseline_creature(n, num_channels=C, radius=R, seed=0)
    K = cr["num_kernels"]
    eng = pkg.make_engine(cr)
    if args.plan:
        vals = [int(v) for v in args.plan.split(",")]
        eng.set_plan(n, *vals)
        plan = eng.plans()[0]
    else:
        plan = eng.plans()[0] if args.no_autotune else eng.autotune(9)[0]
    try:
        import torch
        sync_dev = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is only used for a device-wide barrier
        sync_dev = lambda: None  # noqa: E731

    # The FP64 mass is taken before the warm-up steps (the drift covers warm-up and timed steps),
    # so that nothing but the barrier separates the warm-up from the timed region.
    m0 = eng.mass_fp64()
    eng.run_steps(args.warmup)
    eng.sync()
    sync_dev()
    t0 = time.perf_counter()
    eng.run_steps(args.steps)
    eng.sync()
    sync_dev()
    dt = time.perf_counter() - t0
    m1 = eng.mass_fp64()
    drift = float(abs(m1.sum() - m0.sum()) / m0.sum())
    ms_per_step = dt * 1e3 / args.steps
    value = n * n * args.steps / dt

    # Per-kernel average launch durations: HIP events on the engine stream around one pass per
    # step of 4 * prof_steps direct steps queued back to back (cn_profile_kernels).
    prof_steps = min(args.steps, 50)
    kms = eng.profile_kernels(prof_steps)
    info = eng.info()
    compact = info["compact_kernels"]
    # The boundary's host buffers (mod.rs:216-263): every channel uploaded from host f64
    # (upload_channel, cast to f32) and all channels downloaded as host f32. The timed run's rate
    # with one upload and one download of the state added is reported beside `value`.
    state64 = [np.ascontiguousarray(cr["state"][c], dtype=np.float64).reshape(-1) for c in range(C)]
    io_reps = 5
    t_io = time.perf_counter()
    for _ in range(io_reps):
        for c in range(C):
            eng.upload_channel(state64[c], c)
        eng.sync()
    upload_s = (time.perf_counter() - t_io) / io_reps
    t_io = time.perf_counter()
    for _ in range(io_reps):
        eng.download_all_channels()
    download_s = (time.perf_counter() - t_io) / io_reps
    host_io = {"upload_ms": round(upload_s * 1e3, 3), "download_ms": round(download_s * 1e3, 3),
               "upload_bytes_host": C * n * n * 8, "download_bytes_host": C * n * n * 4,
               "value_with_host_io": round(n * n * args.steps / (dt + upload_s + download_s), 1),
               "note": f"the {args.steps}-step run with one upload (host f64) and one download (host f32) "
                       "of the state added; pageable host buffers"}
    eng.close()
    bpc = algorithmic_bytes_per_cell(C, K, compact, seg_table_bytes_per_cell(info))
    per_kernel = {}
    for name, ms in zip(KERNELS, kms):
        b = bpc[name] * n * n
        per_kernel[name] = {"ms": round(float(ms), 5), "alg_bytes": b, "GBps": round(b / (ms * 1e-3) / 1e9, 1)}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["ms"])
    tag = f"{n}_c{C}k{K}"
    achieved = per_kernel[dom]["GBps"]
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(tag, dom)}
    step_alg = sum(bpc.values()) * n * n
    c2c_model = (36 * C + 32 * K) * n * n  # C2C model: 396 B/cell-step at C3K9
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"DiffusionLenia {n}x{n} C{C}K{K} R={R} periodic, temp 1, seed 0",
                   "grid": n, "channels": C, "kernels": K, "radius": R, "compact_kernels": compact,
                   "column_segments": info["segments"], "segment_radius": info["segment_radius"],
                   "plan": plan},
        "roofline": roof,
        "kernels": per_kernel,
        "step_roofline": {"alg_bytes_per_step": step_alg, "GBps": round(step_alg / (ms_per_step * 1e-3) / 1e9, 1),
                          "frac": round(step_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "c2c_model_bytes_per_step": c2c_model,
                          "c2c_model_frac": round(c2c_model / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "c2c_model_note": "bytes of the full-complex model (36C+32K B/cell) per step time; this "
                                            "design moves fewer bytes, so above 1 means the model is beaten"},
        "mass_drift_rel_fp64": drift,
        "host_io": host_io,
    }
    if not args.no_grad:
        # The second half of the headline metric: batched forward + adjoint creature search
        # against the plain per-creature loop (one engine per creature) and the PyTorch path.
        out["batched_grad"] = batched_grad(pkg, iters=args.iters, torch_iters=args.torch_iters,
                                           plain_iters=args.plain_iters)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
