"""Full-size validation runs of the benchmark configurations (per-step error over 100 steps,
the FP64 error per config, and the ml-rs-example restatement on the same seed).

FP64 reference: the PyTorch restatement of the reference step (oracle/torch_ref.py, itself
pinned to the C++ FP64 oracle to 1e-12 on the CPU) run in float64 on cuda:0, so 100 full-size
FP64 steps take seconds.  Teacher-forced: the FP64 step starts from the engine's previous state.
Bar: <= 1e-5 relative L-inf per step; FP64 mass drift <= 1e-6."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import torch_ref  # noqa: E402

DEV = "cuda:0"
TOL = 1e-5


def _per_step_errors(cn, orc, n, R, steps, seed=0):
    cr = cn.baseline_creature(n, radius=R, seed=seed)
    eng = cn.make_engine(cr)
    kf, src, dst, mu, sg, w = torch_ref.creature_tensors(cr, DEV)
    free = torch.tensor(cr["state"], dtype=torch.float64, device=DEV)
    m0 = eng.mass_fp64().sum()
    errs = []
    prev = cr["state"].astype(np.float64)
    for _ in range(steps):
        eng.iterate()
        got = eng.download_all_channels().reshape(3, n, n)
        want = torch_ref.step(torch.tensor(prev, device=DEV), kf, src, dst, mu, sg, w).cpu().numpy()
        errs.append(orc.rel_linf(got, want))
        free = torch_ref.step(free, kf, src, dst, mu, sg, w)
        prev = got.astype(np.float64)
    free_err = orc.rel_linf(prev, free.cpu().numpy())
    drift = abs(eng.mass_fp64().sum() - m0) / m0
    eng.close()
    return np.array(errs), free_err, drift


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_config1_1024_c3k9_100_steps(cn, orc, seed):
    errs, free_err, drift = _per_step_errors(cn, orc, 1024, 13, 100, seed)
    print(f"1024^2 C3K9 R=13 seed {seed}, 100 steps: per-step rel L-inf vs FP64 max {errs.max():.2e} mean {errs.mean():.2e}; "
          f"free-running after 100 steps {free_err:.2e}; FP64 mass drift {drift:.2e}")
    assert errs.max() <= TOL
    assert drift <= 1e-6


def test_config2_2048_c3k9_per_step(cn, orc):
    errs, free_err, drift = _per_step_errors(cn, orc, 2048, 26, 30)
    print(f"2048^2 C3K9 R=26, 30 steps: per-step rel L-inf vs FP64 max {errs.max():.2e} mean {errs.mean():.2e}; "
          f"free-running after 30 steps {free_err:.2e}; FP64 mass drift {drift:.2e}")
    assert errs.max() <= TOL
    assert drift <= 1e-6


def test_config2_2048_c3k9_1000_steps_sampled(cn, orc):
    """The config's whole 1000-step run (graph replays between the samples): the teacher-forced
    step against FP64 at eight points spread over the run, and the FP64 mass drift of the run."""
    n, R = 2048, 26
    cr = cn.baseline_creature(n, radius=R, seed=0)
    eng = cn.make_engine(cr)
    kf, src, dst, mu, sg, w = torch_ref.creature_tensors(cr, DEV)
    m0 = eng.mass_fp64().sum()
    errs, done = {}, 0
    for s in (1, 2, 10, 100, 250, 500, 750, 1000):
        if s - 1 > done:
            eng.run_steps(s - 1 - done)
        prev = torch.tensor(eng.download_all_channels().reshape(3, n, n), dtype=torch.float64, device=DEV)
        eng.iterate()
        got = torch.tensor(eng.download_all_channels().reshape(3, n, n), dtype=torch.float64, device=DEV)
        want = torch_ref.step(prev, kf, src, dst, mu, sg, w)
        errs[s] = float((got - want).abs().max() / want.abs().max())
        done = s
    drift = abs(eng.mass_fp64().sum() - m0) / m0
    assert eng.info()["steps_done"] == 1000
    eng.close()
    print("2048^2 C3K9 R=26, 1000 steps, teacher-forced step vs FP64 at " +
          ", ".join(f"{s}: {e:.2e}" for s, e in errs.items()) + f"; FP64 mass drift {drift:.2e}")
    assert max(errs.values()) <= TOL
    assert drift <= 1e-6


CHECKPOINTS = (1, 10, 25, 50, 100)
CPU_THREADS = max(1, min(16, os.cpu_count() or 1))


def _three_way(cn, orc, cr, steps=100):
    """The engine, ml-rs-example (oracle/cpu_ref.cpp f32, kernels shared out over threads: the same
    bits as one thread) and the FP64 restatement from the same seed.  Returns the per-step
    teacher-forced engine error and, at CHECKPOINTS, the free-running errors engine vs FP64,
    ml-rs-example vs FP64 and engine vs ml-rs-example; plus the engine's FP64 mass drift."""
    C, n = cr["num_channels"], cr["n"]
    args = (cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"])
    eng = cn.make_engine(cr)
    kf, src, dst, mu, sg, w = torch_ref.creature_tensors(cr, DEV)
    free = torch.tensor(cr["state"], dtype=torch.float64, device=DEV)
    cpu = np.asarray(cr["state"], np.float32)
    m0 = eng.mass_fp64().sum()
    prev = np.asarray(cr["state"], np.float64)
    tf, table, done = [], [], 0
    for t in range(1, steps + 1):
        eng.iterate()
        got = eng.download_all_channels().reshape(C, n, n)
        want = torch_ref.step(torch.tensor(prev, device=DEV), kf, src, dst, mu, sg, w).cpu().numpy()
        tf.append(orc.rel_linf(got, want))
        free = torch_ref.step(free, kf, src, dst, mu, sg, w)
        prev = got.astype(np.float64)
        if t in CHECKPOINTS:
            cpu = orc.run_f32(cpu, *args, steps=t - done, threads=CPU_THREADS)
            done = t
            f64 = free.cpu().numpy()
            table.append((t, orc.rel_linf(got, f64), orc.rel_linf(cpu, f64), orc.rel_linf(got, cpu)))
    drift = abs(eng.mass_fp64().sum() - m0) / m0
    eng.close()
    return np.array(tf), table, drift


def _print_table(title, tf, table, drift):
    print(f"{title}: teacher-forced per-step rel L-inf vs FP64 max {tf.max():.2e} mean {tf.mean():.2e}; "
          f"FP64 mass drift {drift:.2e}")
    print(" step | engine vs FP64 | ml-rs-example vs FP64 | engine vs ml-rs-example (free-running, same seed)")
    for t, a, b, c in table:
        print(f" {t:4d} | {a:14.2e} | {b:21.2e} | {c:.2e}")


def test_config1_against_ml_rs_example_same_seed(cn, orc):
    """The benchmark's ring creature at 1024^2 C3K9 R=13: the engine and the ml-rs-example
    restatement free-running from the same seed for 100 steps, both against FP64."""
    cr = cn.baseline_creature(1024, radius=13, seed=0)
    tf, table, drift = _three_way(cn, orc, cr)
    _print_table("ring creature 1024^2 C3K9 R=13", tf, table, drift)
    assert tf.max() <= TOL and drift <= 1e-6
    for t, e_gpu, e_cpu, e_pair in table:
        # the engine tracks FP64 at least as closely as the reference's own f32 CPU step
        assert e_gpu <= max(2 * e_cpu, t * TOL), (t, e_gpu, e_cpu)
        assert e_pair <= e_gpu + e_cpu + 1e-7, t


def test_example_creature_engine_vs_ml_rs_example(cn, orc):
    """The creature ml-rs-example itself runs (example.rs:310-491; kernels, seed and growth
    parameters from the oracle's f32 restatement of the example's arithmetic, so the engine is
    given exactly the example's inputs) at 1024^2 C3K9 for 100 steps: engine, ml-rs-example and
    FP64 from the same seed."""
    cr = orc.example_creature_f32(1024)
    tf, table, drift = _three_way(cn, orc, cr)
    _print_table("ml-rs-example creature 1024^2 C3K9", tf, table, drift)
    assert tf.max() <= TOL and drift <= 1e-6
    for t, e_gpu, e_cpu, e_pair in table:
        assert e_gpu <= max(2 * e_cpu, t * TOL), (t, e_gpu, e_cpu)
        assert e_pair <= e_gpu + e_cpu + 1e-7, t
