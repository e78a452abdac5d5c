"""Audio consumer of the rendered frames: the reference's sonar-ping synthesiser
(ml-rs/audio.rs) over the C ABI.

`AudioSynth` mirrors the reference's `AudioSynth` (audio.rs:104-396): `extract_features` of an
RGB24 frame (or `GpuMaceLenia.render_features()`, the same features reduced on the device
without downloading the frame), `generate_frame` for one video frame's samples, and
`activity_level`; `write_wav` is audio.rs:404-441.  The synthesis is host DSP in the library
(csrc/audio.cpp); nothing here computes samples.
"""
import ctypes

import numpy as np

from ._lib import check, lib

SAMPLE_RATE = 44100  # main.rs:220


class AudioSynth:
    def __init__(self, sample_rate=SAMPLE_RATE):
        h = ctypes.c_void_p()
        check(lib().cn_audio_create(int(sample_rate), ctypes.byref(h)))
        self._h = h
        self.sample_rate = int(sample_rate)

    def close(self):
        if self._h:
            lib().cn_audio_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def extract_features(render_rgb, grid_size):
        """(means, variances) of RGB24 bytes / 255 per channel (audio.rs:181-209)."""
        rgb = np.ascontiguousarray(render_rgb, dtype=np.uint8).reshape(-1)
        if rgb.size != 3 * grid_size * grid_size:
            raise ValueError(f"expected {3 * grid_size * grid_size} bytes, got {rgb.size}")
        means = np.empty(3, dtype=np.float32)
        variances = np.empty(3, dtype=np.float32)
        check(lib().cn_audio_extract_features(rgb.ctypes.data, int(grid_size), means.ctypes.data,
                                              variances.ctypes.data))
        return means, variances

    def generate_frame(self, features, num_samples):
        """Interleaved stereo int16 samples (2 * num_samples) for one video frame."""
        means = np.ascontiguousarray(features[0], dtype=np.float32)
        variances = np.ascontiguousarray(features[1], dtype=np.float32)
        if means.size != 3 or variances.size != 3:
            raise ValueError("features are three means and three variances")
        out = np.empty(2 * int(num_samples), dtype=np.int16)
        check(lib().cn_audio_generate_frame(self._h, means.ctypes.data, variances.ctypes.data, int(num_samples),
                                            out.ctypes.data, out.size))
        return out

    def activity_level(self):
        v = ctypes.c_float()
        check(lib().cn_audio_activity(self._h, ctypes.byref(v)))
        return v.value


def write_wav(path, samples, sample_rate=SAMPLE_RATE, num_channels=2):
    """16-bit PCM WAV of interleaved samples (audio.rs:404-441)."""
    s = np.ascontiguousarray(samples, dtype=np.int16).reshape(-1)
    check(lib().cn_write_wav(str(path).encode(), s.ctypes.data, s.size, int(sample_rate), int(num_channels)))
