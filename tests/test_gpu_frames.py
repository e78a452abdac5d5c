"""GPU end-to-end frame run in the reference's layout (save_pngs.rs:66-168): creature files
written by the writers, frames saved by the harness command, compared (compare_pngs.py:30-80)
with frames of the FP64 oracle's free-running states: every frame within 1 gray level, every
state within 1e-5 relative L-inf per step."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_frame_run_matches_oracle_frames(cn, orc, tmp_path, monkeypatch):
    n, steps = 64, 5
    cr = cn.baseline_creature(n, radius=9, seed=11, patch=32)
    os.makedirs(tmp_path / "seed")
    os.makedirs(tmp_path / "kernels")
    # The JSON convention: JSON c0 is the output channel, c1 the input (config.rs:40).
    cfg = cn.SimConfig(seed_size=n, num_channels=3, num_kernels=9,
                       seed_channels=[cr["state"][c].reshape(-1).astype(np.float64) for c in range(3)],
                       c0=cr["src"], c1=cr["dst"], growth_mu=cr["mu"], growth_sigma=cr["sigma"],
                       growth_weights=cr["weights"])
    cn.write_config(tmp_path / "seed" / "blob.json", cfg)
    cn.write_kernels(tmp_path / "kernels" / f"blob_{n}.bin", cr["kernels_fft"])
    monkeypatch.chdir(tmp_path)
    assert cn.frames._main(["--creature", "blob", "--steps", str(steps), "--out", "pngs", "--rgb"]) == 0

    st = cr["state"].astype(np.float64)
    for step in range(steps + 1):
        if step:
            st = orc.run_f64(st, cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"])
        cn.frames.write_png(f"pngs/ref_{step:04d}.png", cn.frames.gray_frame(st.astype(np.float32)))
        np.save(f"pngs/ref_{step:04d}.npy", st)
    rows = cn.frames.compare_frames("pngs/cn_frame_", "pngs/ref_")
    assert len(rows) == steps + 1
    for r in rows:
        print(r)
        assert r["ok"], r
        assert r["rel_linf"] <= 1e-5 * max(1, r["step"]), r
    rgb = cn.frames.read_png(f"pngs/cn_frame_rgb_{steps:04d}.png")
    last = np.load(f"pngs/cn_frame_{steps:04d}.npy")
    assert np.array_equal(rgb, orc.render_rgb(last))
