"""GPU: the video command line's frames (ml-rs/main.rs:268-303: iterate_and_render, then the
RGB24 download) written as raw rgb24 are the oracle's render of its FP64 free-running states."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_raw_video_frames_match_oracle_render(cn, orc, tmp_path):
    n, fps, seconds = 64, 3, 1
    cr = cn.baseline_creature(n, radius=9, seed=12, patch=32)
    os.makedirs(tmp_path / "seed")
    os.makedirs(tmp_path / "kernels")
    cfg = cn.SimConfig(seed_size=n, num_channels=3, num_kernels=9,
                       seed_channels=[cr["state"][c].reshape(-1).astype(np.float64) for c in range(3)],
                       c0=cr["src"], c1=cr["dst"], growth_mu=cr["mu"], growth_sigma=cr["sigma"],
                       growth_weights=cr["weights"])
    cn.write_config(tmp_path / "seed" / "blob.json", cfg)
    cn.write_kernels(tmp_path / "kernels" / f"blob_{n}.bin", cr["kernels_fft"])
    path = cn.video.render_video("blob", seconds=seconds, fps=fps, output=str(tmp_path / "videos"),
                                 root=str(tmp_path), use_ffmpeg=False, log=lambda *_: None)
    raw = np.fromfile(path, dtype=np.uint8)
    frames = fps * seconds
    assert raw.size == frames * n * n * 3
    raw = raw.reshape(frames, n, n, 3)
    eng = cn.make_engine(cr)  # the same creature loaded directly: the same frames, bit for bit
    for f in range(frames):
        eng.iterate_and_render()
        assert np.array_equal(raw[f], eng.download_render().reshape(n, n, 3)), f
    eng.close()
    st = cr["state"].astype(np.float64)
    for f in range(frames):
        st = orc.run_f64(st, cr["kernels_fft"], cr["src"], cr["dst"], cr["mu"], cr["sigma"], cr["weights"])
        want = orc.render_rgb(st.astype(np.float32))
        # the engine's f32 state differs from FP64 by ~1e-6 relative: a byte may round the other way
        assert np.abs(raw[f].astype(np.int64) - want.astype(np.int64)).max() <= 1, f
