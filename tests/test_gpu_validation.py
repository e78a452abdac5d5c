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


def test_config1_against_ml_rs_example_same_seed(cn, orc):
    """The engine and the ml-rs-example restatement (oracle/cpu_ref.cpp, f32) free-running from the
    same seed for 3 steps; both against the FP64 oracle."""
    n, steps = 1024, 3
    cr = cn.baseline_creature(n, radius=13, seed=0)
    args = (cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"])
    eng = cn.make_engine(cr)
    eng.run_steps(steps)
    gpu = eng.download_all_channels().reshape(3, n, n)
    eng.close()
    cpu = orc.run_f32(cr["state"], *args, steps=steps)
    kf, src, dst, mu, sg, w = torch_ref.creature_tensors(cr, DEV)
    fp64 = torch_ref.rollout(torch.tensor(cr["state"], dtype=torch.float64, device=DEV), kf, src, dst, mu, sg, w,
                             steps).cpu().numpy()
    e_gpu, e_cpu, e_pair = orc.rel_linf(gpu, fp64), orc.rel_linf(cpu, fp64), orc.rel_linf(gpu, cpu)
    print(f"1024^2 C3K9, {steps} steps from the same seed: engine vs FP64 {e_gpu:.2e}, ml-rs-example (f32) vs FP64 "
          f"{e_cpu:.2e}, engine vs ml-rs-example {e_pair:.2e}")
    assert e_gpu <= steps * TOL
    assert e_pair <= steps * TOL + e_cpu
