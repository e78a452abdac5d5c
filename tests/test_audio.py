"""Audio consumer (ml-rs/audio.rs) on the CPU.

The oracle (oracle/audio_ref.py) restates audio.rs; each of its stages is pinned here by a
closed form or an independent computation (numpy statistics, the LCG's closed form, the one-pole
filter's impulse and step responses, the ping-pong positions of the delay line, Python's `wave`
module for the file format).  The library's synthesiser (csrc/audio.cpp) must then produce the
oracle's samples exactly, and its features within one f32 rounding of the reference's f64 sums.
"""
import math
import os
import sys
import wave

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def aud():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import audio_ref
    finally:
        sys.path.pop(0)
    return audio_ref


@pytest.fixture(scope="module")
def au(cn):
    from cella_nova_amd import audio
    return audio


# --- the oracle, pinned -------------------------------------------------------------------------

def test_lcg_matches_its_closed_form(aud):
    # x_k = a^k x_0 + c (a^k - 1) / (a - 1)  (mod 2^32), evaluated with exact integers.
    a, c, m, x0 = 1103515245, 12345, 1 << 32, 0xC0FFEE
    x = x0
    for k in range(1, 200):
        x = aud.lcg_next(x)
        ak = pow(a, k)
        assert x == (ak * x0 + c * ((ak - 1) // (a - 1))) % m, k
    assert aud.lcg_unit(0) == 0.0 and aud.lcg_unit(0xFFFFFFFF) == 1.0


def test_features_against_numpy(aud):
    rng = np.random.default_rng(7)
    for grid in (4, 16, 64):
        rgb = rng.integers(0, 256, size=(grid, grid, 3), dtype=np.uint8)
        means, variances = aud.extract_features(rgb, grid)
        v = rgb.reshape(-1, 3).astype(np.float64) / 255.0
        assert np.allclose(means, v.mean(axis=0), rtol=0, atol=1e-7)
        assert np.allclose(variances, v.var(axis=0), rtol=0, atol=2e-7)
    # closed forms: black, white (variance clamps at 0), one white pixel in black
    z = np.zeros((8, 8, 3), np.uint8)
    assert aud.extract_features(z, 8)[0].tolist() == [0, 0, 0]
    m, var = aud.extract_features(np.full((8, 8, 3), 255, np.uint8), 8)
    assert m.tolist() == [1, 1, 1] and var.tolist() == [0, 0, 0]
    z[3, 5, 1] = 255
    m, var = aud.extract_features(z, 8)
    assert m[1] == np.float32(1 / 64) and abs(var[1] - (1 / 64 - 1 / 64 ** 2)) < 1e-8


def test_envelope_closed_form(aud):
    assert aud.envelope(1000, 1000) == 0.0             # attack starts at 0
    assert abs(aud.envelope(950, 1000) - 0.7) < 1e-6    # (1 - 0.95) * 14
    assert aud.envelope(930, 1000) == 1.0               # progress 0.93: the decay branch, (1)^1.6
    assert abs(aud.envelope(500, 1000) - (0.5 / 0.93) ** 1.6) < 1e-6
    assert aud.envelope(0, 1000) == 0.0


def test_lowpass_impulse_and_step_response(aud):
    sr = 44100
    alpha = aud.lowpass_alpha(sr)
    assert abs(alpha - (1 - math.exp(-2 * math.pi * 1400 / sr))) < 1e-7
    s = np.float32(0)
    for n in range(300):  # impulse: y_n = alpha (1 - alpha)^n
        s = aud.lowpass_step(s, 1.0 if n == 0 else 0.0, alpha)
        assert abs(s - alpha * (1 - alpha) ** n) <= 1e-6 * alpha, n
    s = np.float32(0)
    for n in range(300):  # step: y_n = 1 - (1 - alpha)^(n + 1)
        s = aud.lowpass_step(s, 1.0, alpha)
        assert abs(s - (1 - (1 - alpha) ** (n + 1))) <= 1e-6, n


def test_delay_line_ping_pongs(aud):
    # An impulse on the left at frame 0 returns on the right after 3307 frames and on the left
    # 3308 frames later: the 6615-sample delay of an interleaved buffer crosses the channels.
    # Echo k carries 0.30 * 0.35^(k - 1).
    buf = np.zeros(aud.DELAY_BUFFER_SIZE, np.float32)
    pos, hits = 0, []
    for j in range(14000):
        wl, wr, pos = aud.delay_step(buf, pos, 1.0 if j == 0 else 0.0, 0.0)
        if wl:
            hits.append((j, "L", float(wl)))
        if wr:
            hits.append((j, "R", float(wr)))
    want = [(0, "L", 1.0), (3307, "R", 0.3), (6615, "L", 0.3 * 0.35), (9922, "R", 0.3 * 0.35 ** 2),
            (13230, "L", 0.3 * 0.35 ** 3)]
    assert [(j, c) for j, c, _ in hits] == [(j, c) for j, c, _ in want]
    for (_, _, got), (_, _, w) in zip(hits, want):
        assert abs(got - w) < 1e-6 * max(w, 1e-3)


def test_limiter(aud):
    assert aud.to_i16(0.0) == 0
    assert aud.to_i16(10.0) == 32767 and aud.to_i16(-10.0) == -32767
    for x in np.linspace(-0.5, 0.5, 101):
        assert abs(aud.to_i16(x) - math.trunc(math.tanh(5 * x) * 32767)) <= 1


def test_silence_without_variance_change(aud):
    s = aud.Synth(44100)
    f = (np.float32([0.3, 0.4, 0.5]), np.float32([0.01, 0.02, 0.03]))
    for _ in range(4):
        assert not s.generate_frame(*f, 200).any()
    assert s.activity == 0.0


def test_single_ping_pan_and_duration(aud):
    # One variance change of 1e-5 on channel 1: a ping of 250 + 2750 * 0.2 = 800 samples, amplitude
    # 0.234.  Before the first echo the filtered left/right signals stay in the ratio
    # (1 - pan) / pan (same linear filter on both), with pan in 0.5 +- 0.2.
    s = aud.Synth(44100)
    m = np.float32([0.2, 0.3, 0.4])
    v0 = np.float32([0.01, 0.02, 0.03])
    s.generate_frame(m, v0, 10)
    v1 = v0.copy()
    v1[1] = np.float32(v0[1] + np.float32(1e-5))
    out = s.generate_frame(m, v1, 3000).astype(np.float64).reshape(-1, 2)
    assert len(s.pings) == 0  # the ping has ended within the frame
    pre = np.arctanh(np.clip(out / 32767.0, -0.999, 0.999)) / 5.0
    live = np.abs(pre[:, 1]) > 0.02
    ratio = np.median(pre[live, 0] / pre[live, 1])
    pan = 1 / (1 + ratio)
    assert 0.3 <= pan <= 0.7
    assert s.cooldown == [0, 2 + int(8 * 0.8), 0]


def test_wav_writer_against_python_wave(au, tmp_path):
    rng = np.random.default_rng(3)
    samples = rng.integers(-32768, 32768, size=2 * 1000, dtype=np.int16)
    p = tmp_path / "a.wav"
    au.write_wav(p, samples, 44100, 2)
    with wave.open(str(p), "rb") as w:
        assert (w.getnchannels(), w.getsampwidth(), w.getframerate(), w.getnframes()) == (2, 2, 44100, 1000)
        assert w.readframes(1000) == samples.astype("<i2").tobytes()
    q = tmp_path / "b.wav"
    with wave.open(str(q), "wb") as w:
        w.setnchannels(2)
        w.setsampwidth(2)
        w.setframerate(44100)
        w.writeframes(samples.astype("<i2").tobytes())
    assert p.read_bytes() == q.read_bytes()


# --- the library against the oracle -----------------------------------------------------------

def test_features_match_the_oracle(au, aud):
    rng = np.random.default_rng(11)
    for grid in (16, 128):
        for kind in ("uniform", "sparse"):
            rgb = rng.integers(0, 256, size=(grid, grid, 3), dtype=np.uint8)
            if kind == "sparse":
                rgb[rng.random((grid, grid)) < 0.9] = 0
            m, v = au.AudioSynth.extract_features(rgb, grid)
            mo, vo = aud.extract_features(rgb, grid)
            # exact integer sums against the reference's f64 running sums: one f32 rounding apart
            assert np.all(np.abs(m - mo) <= np.spacing(np.maximum(mo, np.float32(1e-30)))), (m, mo)
            assert np.all(np.abs(v - vo) <= 2e-7), (v, vo)


def _scripted_features(frames, seed, lo=-7.0):
    """Per-frame features whose variances jump often by 10^U(lo, -4) (pings on every channel,
    cooldowns, and, for large jumps, more than ten live pings at once)."""
    rng = np.random.default_rng(seed)
    m = rng.random(3).astype(np.float32)
    v = (rng.random(3) * 0.05).astype(np.float32)
    out = []
    for f in range(frames):
        if f % 7 != 3:  # some frames change nothing
            step = np.float32(10) ** rng.uniform(lo, -4, size=3).astype(np.float32)
            v = (v + step * rng.choice([-1, 1], size=3)).astype(np.float32)
            m = np.clip(m + rng.normal(0, 0.01, 3), 0, 1).astype(np.float32)
        out.append((m.copy(), np.maximum(v, 0).astype(np.float32)))
    return out


@pytest.mark.parametrize("sr,per_frame,frames,lo", [(44100, 100, 70, -4.5), (44100, 735, 12, -7.0),
                                                   (8000, 133, 30, -7.0)])
def test_synth_matches_the_oracle_sample_for_sample(au, aud, sr, per_frame, frames, lo):
    feats = _scripted_features(frames, seed=sr + per_frame, lo=lo)
    lib_s = au.AudioSynth(sr)
    ora = aud.Synth(sr)
    most = 0
    for f, (m, v) in enumerate(feats):
        n = per_frame + (1 if f < 3 else 0)
        got = lib_s.generate_frame((m, v), n)
        want = ora.generate_frame(m, v, n)
        assert np.array_equal(got, want), (f, np.flatnonzero(got != want)[:5])
        assert lib_s.activity_level() == ora.activity
        most = max(most, len(ora.pings))
    lib_s.close()
    if per_frame == 100:
        assert most == aud.MAX_PINGS  # the polyphony limit was reached (quietest ping dropped)


def test_argument_errors(au, cn):
    from cella_nova_amd._lib import CellaNovaError
    with pytest.raises(CellaNovaError):
        au.AudioSynth(0)
    with pytest.raises(ValueError):
        au.AudioSynth.extract_features(np.zeros(10, np.uint8), 4)
    s = au.AudioSynth()
    with pytest.raises(ValueError):
        s.generate_frame((np.zeros(2, np.float32), np.zeros(3, np.float32)), 10)
    assert s.generate_frame((np.zeros(3, np.float32), np.zeros(3, np.float32)), 0).size == 0
