"""Oracle (test infrastructure): a CPU restatement of the reference's audio consumer,
ml-rs/audio.rs, used only by the tests as the checker of the library's synthesiser
(csrc/audio.cpp) and device frame statistics.  Nothing in the product imports it.

Arithmetic is f32 operation by operation as in audio.rs (numpy float32 scalars: one IEEE
rounding per operation, no fused multiply-adds); sin/exp/pow/tanh/fmod are the C library's f32
functions (Rust's f32 methods call the same libm), reached through ctypes.  Each stage is a
function of its own so that the tests can pin it by a closed form.
"""
import ctypes
import ctypes.util
import math

import numpy as np

F = np.float32
_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
for _name in ("sinf", "expf", "tanhf"):
    getattr(_libm, _name).restype = ctypes.c_float
    getattr(_libm, _name).argtypes = [ctypes.c_float]
for _name in ("powf", "fmodf"):
    getattr(_libm, _name).restype = ctypes.c_float
    getattr(_libm, _name).argtypes = [ctypes.c_float, ctypes.c_float]


def sinf(x):
    return F(_libm.sinf(float(x)))


def expf(x):
    return F(_libm.expf(float(x)))


def tanhf(x):
    return F(_libm.tanhf(float(x)))


def powf(x, y):
    return F(_libm.powf(float(x), float(y)))


def fmodf(x, y):
    return F(_libm.fmodf(float(x), float(y)))


TAU = F(2 * math.pi)               # std::f32::consts::TAU
MAX_PINGS = 10                     # audio.rs:22
PING_MIN_SAMPLES = 250             # audio.rs:25
PING_MAX_SAMPLES = 3000            # audio.rs:26
COOLDOWN_MIN, COOLDOWN_MAX = 2, 10  # audio.rs:29-30
VAR_CHANGE_THRESHOLD = F(0.000001)  # audio.rs:33
DELAY_BUFFER_SIZE = 16384          # audio.rs:36
DELAY_TIME_SAMPLES = 6615          # audio.rs:37
DELAY_FEEDBACK = F(0.35)           # audio.rs:38
DELAY_WET = F(0.30)                # audio.rs:39
LP_CUTOFF = F(1400.0)              # audio.rs:42
TREMOLO_RATE = F(6.0)              # audio.rs:45
BASE_FREQ = (F(100.0), F(300.0), F(750.0))  # audio.rs:136-150
BASE_PAN = (F(0.15), F(0.5), F(0.85))       # audio.rs:287


def extract_features(rgb, grid):
    """audio.rs:181-209: sequential f64 sums of byte / 255 and its square per channel (left to
    right over the pixels, as np.add.accumulate adds), then f32 mean and variance."""
    px = np.asarray(rgb, dtype=np.uint8).reshape(grid * grid, 3)
    n_f = float(grid * grid)
    means = np.zeros(3, np.float32)
    variances = np.zeros(3, np.float32)
    for ch in range(3):
        v = px[:, ch].astype(np.float64) / 255.0
        s = float(np.add.accumulate(v)[-1])
        sq = float(np.add.accumulate(v * v)[-1])
        means[ch] = F(s / n_f)
        mean_sq = F(sq / n_f)
        variances[ch] = max(F(mean_sq - F(means[ch] * means[ch])), F(0.0))
    return means, variances


def lcg_next(seed):
    """audio.rs:168-174: seed * 1103515245 + 12345 mod 2^32."""
    return (seed * 1103515245 + 12345) & 0xFFFFFFFF


def lcg_unit(seed):
    """The f32 in [0, 1] a seed maps to: seed as f32 / u32::MAX as f32."""
    return F(F(seed) / F(4294967295))


def envelope(remaining, total):
    """audio.rs:343-349: linear attack over the first 7 %, then (progress / 0.93)^1.6."""
    progress = F(F(remaining) / F(total))
    env = F(F(1.0) - progress) * F(14.0) if progress > F(0.93) else powf(F(progress / F(0.93)), F(1.6))
    return min(max(F(env), F(0.0)), F(1.0))


def lowpass_alpha(sample_rate):
    """audio.rs:316: 1 - exp(-TAU * cutoff / sr)."""
    return F(F(1.0) - expf(F(F(-TAU * LP_CUTOFF) / F(sample_rate))))


def lowpass_step(state, x, alpha):
    """audio.rs:371-372: one-pole low-pass, state += alpha * (x - state)."""
    return F(state + F(alpha * F(x - state)))


def delay_step(buf, pos, fl, fr):
    """audio.rs:377-389: interleaved feedback delay line; returns (wet_l, wet_r, new pos)."""
    read = (pos + DELAY_BUFFER_SIZE - DELAY_TIME_SAMPLES) % DELAY_BUFFER_SIZE
    el = buf[read]
    er = buf[(read + 1) % DELAY_BUFFER_SIZE]
    wet_l = F(fl + F(el * DELAY_WET))
    wet_r = F(fr + F(er * DELAY_WET))
    buf[pos] = F(fl + F(el * DELAY_FEEDBACK))
    buf[(pos + 1) % DELAY_BUFFER_SIZE] = F(fr + F(er * DELAY_FEEDBACK))
    return wet_l, wet_r, (pos + 2) % DELAY_BUFFER_SIZE


def to_i16(x):
    """audio.rs:391-395: tanh(5 x) * 32767, clamped, truncated toward zero (Rust `as i16`)."""
    y = F(tanhf(F(x * F(5.0))) * F(32767.0))
    if math.isnan(y):
        return 0
    y = min(max(float(y), -32768.0), 32767.0)
    return int(y)


class Synth:
    """AudioSynth (audio.rs:104-396)."""

    def __init__(self, sample_rate):
        self.sr = sample_rate
        self.pings = []  # [phase, f0, f1, remaining, total, amplitude, pan]
        self.cooldown = [0, 0, 0]
        self.prev_var = [F(0.0)] * 3
        self.has_prev = False
        self.buf = np.zeros(DELAY_BUFFER_SIZE, np.float32)
        self.pos = 0
        self.lp_l = F(0.0)
        self.lp_r = F(0.0)
        self.trem_phase = F(0.0)
        self.seed = 0xC0FFEE
        self.activity = F(0.0)

    def rand(self):
        self.seed = lcg_next(self.seed)
        return lcg_unit(self.seed)

    def rand_range(self, lo, hi):
        return F(F(lo) + F(self.rand() * F(F(hi) - F(lo))))

    def fire_ping(self, ch, change, means):  # audio.rs:245-306
        if len(self.pings) >= MAX_PINGS:
            amps = [p[5] for p in self.pings]
            self.pings.pop(amps.index(min(amps)))
        center = F(BASE_FREQ[ch] * F(F(0.8) + F(F(means[ch]) * F(0.5))))
        if self.rand() > F(0.5):
            f0, f1 = F(center * F(0.6)), F(center * F(1.4))
        else:
            f0, f1 = F(center * F(1.5)), F(center * F(0.5))
        change_norm = min(F(change / F(0.00005)), F(1.0))
        duration = PING_MIN_SAMPLES + int(F(F(PING_MAX_SAMPLES - PING_MIN_SAMPLES) * change_norm))
        amplitude = max(min(F(change * F(23400.0)), F(0.95)), F(0.10))
        pan = min(max(F(BASE_PAN[ch] + self.rand_range(-0.2, 0.2)), F(0.0)), F(1.0))
        self.cooldown[ch] = COOLDOWN_MIN + int(F(F(COOLDOWN_MAX - COOLDOWN_MIN) * F(F(1.0) - change_norm)))
        phase = self.rand_range(0.0, TAU)
        self.pings.append([phase, f0, f1, duration, duration, amplitude, pan])

    def trigger(self, means, variances):  # audio.rs:213-243
        if not self.has_prev:
            self.prev_var = [F(v) for v in variances]
            self.has_prev = True
            return
        total = F(0.0)
        for ch in range(3):
            if self.cooldown[ch] > 0:
                self.cooldown[ch] -= 1
            change = F(abs(F(F(variances[ch]) - self.prev_var[ch])))
            total = F(total + change)
            self.prev_var[ch] = F(variances[ch])
            if self.cooldown[ch] > 0 or change < VAR_CHANGE_THRESHOLD:
                continue
            self.fire_ping(ch, change, means)
        self.activity = F(F(self.activity * F(0.85)) + F(total * F(0.15)))

    def generate_frame(self, means, variances, num_samples):  # audio.rs:310-396
        self.trigger(means, variances)
        sr = F(self.sr)
        alpha = lowpass_alpha(self.sr)
        out = np.zeros(2 * num_samples, np.int16)
        for i in range(num_samples):
            t = F(i)
            trem = F(F(sinf(F(F(F(F(TAU * TREMOLO_RATE) * t) / sr) + self.trem_phase)) * F(0.15)) + F(0.85))
            left = F(0.0)
            right = F(0.0)
            for p in self.pings:
                if p[3] == 0:
                    continue
                env = envelope(p[3], p[4])
                tf = F(F(1.0) - F(F(p[3]) / F(p[4])))
                freq = F(p[1] + F(F(p[2] - p[1]) * tf))
                inc = F(F(TAU * freq) / sr)
                tone = sinf(p[0])
                p[0] = F(p[0] + inc)
                if p[0] >= TAU:
                    p[0] = F(p[0] - TAU)
                harm2 = F(F(sinf(F(p[0] * F(2.0))) * F(0.15)) * env)
                sample = F(F(F(F(tone + harm2) * env) * p[5]) * trem)
                left = F(left + F(sample * F(F(1.0) - p[6])))
                right = F(right + F(sample * p[6]))
                p[3] = max(p[3] - 1, 0)
            self.pings = [p for p in self.pings if p[3] > 0]
            self.lp_l = lowpass_step(self.lp_l, left, alpha)
            self.lp_r = lowpass_step(self.lp_r, right, alpha)
            wl, wr, self.pos = delay_step(self.buf, self.pos, self.lp_l, self.lp_r)
            out[2 * i] = to_i16(wl)
            out[2 * i + 1] = to_i16(wr)
        self.trem_phase = F(self.trem_phase + F(F(F(TAU * TREMOLO_RATE) * F(num_samples)) / sr))
        self.trem_phase = fmodf(self.trem_phase, TAU)
        return out
