"""GPU: the audio features of rendered frames reduced on the device (cn_creature_render_features)
against the same features of the downloaded RGB24 bytes (exact: both are integer byte sums) and
against the oracle's restatement of audio.rs:181-209 (f64 running sums, one f32 rounding
apart); the video command line with --with-sound writes the synthesiser's samples for the
device features of every frame."""
import os
import sys
import wave

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def aud():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import audio_ref
    finally:
        sys.path.pop(0)
    return audio_ref


def _check(au, aud, dev, rgb, n):
    m, v = dev
    mh, vh = au.AudioSynth.extract_features(rgb, n)
    assert np.array_equal(m, mh) and np.array_equal(v, vh)
    mo, vo = aud.extract_features(rgb, n)
    assert np.all(np.abs(m - mo) <= np.spacing(np.maximum(mo, np.float32(1e-30))))
    assert np.all(np.abs(v - vo) <= 2e-7)


@pytest.mark.parametrize("n", [64, 256, 2048])
def test_device_features_match_host_and_oracle(cn, aud, n):
    from cella_nova_amd import audio as au
    cr = cn.baseline_creature(n, radius=max(4, n // 80), seed=5, patch=min(n, 256))
    eng = cn.make_engine(cr)
    m, v = eng.render_features()  # before any rendered frame: a black frame
    assert not m.any() and not v.any()
    for _ in range(3):
        eng.iterate_and_render()
        _check(au, aud, eng.render_features(), eng.download_render().reshape(n, n, 3), n)
    eng.close()


def test_batch_features_per_creature(cn, aud):
    from cella_nova_amd import audio as au
    specs = [cn.baseline_creature(64, radius=6, seed=1, patch=32),
             cn.baseline_creature(128, num_channels=2, radius=8, seed=2, patch=64),
             cn.baseline_creature(64, num_channels=4, radius=6, seed=3, patch=32)]
    b = cn.CreatureBatch(specs)
    for _ in range(2):
        b.iterate_and_render()
    for i, cr in enumerate(specs):
        n = cr["state"].shape[-1]
        _check(au, aud, b.creature_render_features(i), b.creature_download_render(i), n)
    b.close()


def test_video_with_sound(cn, tmp_path):
    from cella_nova_amd import audio as au
    n, fps, seconds = 64, 7, 2
    cr = cn.baseline_creature(n, radius=9, seed=12, patch=32)
    os.makedirs(tmp_path / "seed")
    os.makedirs(tmp_path / "kernels")
    cfg = cn.SimConfig(seed_size=n, num_channels=3, num_kernels=9,
                       seed_channels=[cr["state"][c].reshape(-1).astype(np.float64) for c in range(3)],
                       c0=cr["src"], c1=cr["dst"], growth_mu=cr["mu"], growth_sigma=cr["sigma"],
                       growth_weights=cr["weights"])
    cn.write_config(tmp_path / "seed" / "blob.json", cfg)
    cn.write_kernels(tmp_path / "kernels" / f"blob_{n}.bin", cr["kernels_fft"])
    out = tmp_path / "videos"
    cn.video.render_video("blob", seconds=seconds, fps=fps, output=str(out), root=str(tmp_path),
                          use_ffmpeg=False, log=lambda *_: None, with_sound=True)
    with wave.open(str(out / "blob.wav"), "rb") as w:
        assert (w.getnchannels(), w.getframerate(), w.getnframes()) == (2, 44100, seconds * 44100)
        got = np.frombuffer(w.readframes(w.getnframes()), dtype="<i2")
    # The same creature replayed directly: host features of each downloaded frame into a fresh
    # synthesiser, with main.rs's split of 88200 samples over 14 frames (6300 each).
    eng = cn.make_engine(cr)
    syn = au.AudioSynth()
    total = fps * seconds
    per, rem = divmod(seconds * 44100, total)
    want = []
    for f in range(total):
        eng.iterate_and_render()
        feats = au.AudioSynth.extract_features(eng.download_render(), n)
        want.append(syn.generate_frame(feats, per + (1 if f < rem else 0)))
    eng.close()
    assert np.array_equal(got, np.concatenate(want))
    assert np.abs(got).max() > 0  # the creature's changes produced pings
