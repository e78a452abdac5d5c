"""Video command line of the reference (ml-rs/main.rs:170-365), with its audio consumer.

For every creature (--creature NAME, or every seed/*.json as main.rs:128-168 discovers them): load
seed/NAME.json and kernels/NAME_<size>.bin with size = seed_size rounded up to a power of two
(main.rs:80-125), then seconds * fps frames of `iterate_and_render` + `download_render`
(main.rs:268-303): RGB24 frames of size x size.  The frames go to ffmpeg's stdin with the
reference's encoder arguments (main.rs:238-264) when an `ffmpeg` executable is on PATH, otherwise
to <output>/<NAME>.rgb (raw rgb24, the same bytes ffmpeg would have read; the command that
encodes it is printed).  An existing <output>/<NAME>.mp4 is skipped as in main.rs:203-207.
--with-sound (main.rs:219-235, 286-296, 315-361): one AudioSynth per video at 44.1 kHz, fed per
frame with the frame's features (reduced on the device from the packed frame instead of the
downloaded bytes; the same values) and seconds * 44100 / frames samples, the remainder spread
over the first frames; the samples go to <output>/<NAME>.wav, which ffmpeg muxes into the video
(AAC, -shortest) and which is then removed; without ffmpeg the WAV stays beside the raw frames.
"""
import argparse
import os
import shutil
import subprocess
import sys


def discover_creatures(root="."):
    """Names of seed/*.json, sorted (main.rs:128-168)."""
    d = os.path.join(root, "seed")
    if not os.path.isdir(d):
        return []
    return sorted(f[:-5] for f in os.listdir(d) if f.endswith(".json"))


def setup_simulation(name, temp=1.0, root="."):
    """(engine, size): a GpuMaceLenia loaded with seed/NAME.json and kernels/NAME_<size>.bin
    (main.rs:80-125; the c0/c1 swap happens in load_config, config.rs:40)."""
    from .config import load_config, load_kernels
    from .engine import GpuMaceLenia
    cfg = load_config(os.path.join(root, "seed", f"{name}.json"))
    size = 1 << max(0, (int(cfg.seed_size) - 1).bit_length())
    kers = load_kernels(os.path.join(root, "kernels", f"{name}_{size}.bin"), cfg.num_kernels, size)
    eng = GpuMaceLenia(size, cfg.num_channels, cfg.num_kernels, cfg.c0, cfg.c1, cfg.growth_mu, cfg.growth_sigma,
                       cfg.growth_weights, temp=temp)
    for k in range(cfg.num_kernels):
        eng.set_kernel(kers[k], k)
    for c in range(cfg.num_channels):
        eng.upload_channel(cfg.seed_channels[c], c)
    return eng, size


def ffmpeg_args(size, fps, path):
    """The reference's encoder invocation for raw rgb24 frames on stdin (main.rs:238-260)."""
    return ["ffmpeg", "-y", "-f", "rawvideo", "-pixel_format", "rgb24", "-video_size", f"{size}x{size}",
            "-framerate", str(fps), "-i", "-", "-c:v", "libx264", "-crf", "18", "-preset", "medium",
            "-pix_fmt", "yuv420p", path]


def mux_args(video, wav, out):
    """The reference's audio mux invocation (main.rs:329-345)."""
    return ["ffmpeg", "-y", "-i", video, "-i", wav, "-c:v", "copy", "-c:a", "aac", "-shortest", out]


def render_video(name, seconds=60, fps=60, output="videos", temp=1.0, root=".", use_ffmpeg=None, log=print,
                 with_sound=False):
    """Renders one creature; returns the path written (None when the video already exists)."""
    mp4 = os.path.join(output, f"{name}.mp4")
    if os.path.exists(mp4):
        log("  Video already exists, skipping.")
        return None
    os.makedirs(output, exist_ok=True)
    if use_ffmpeg is None:
        use_ffmpeg = shutil.which("ffmpeg") is not None
    eng, size = setup_simulation(name, temp, root)
    total = int(seconds) * int(fps)
    log(f"Generating {total} frames at {fps} FPS ({seconds} seconds)")
    synth, chunks = None, []
    if with_sound:
        from .audio import SAMPLE_RATE, AudioSynth
        synth = AudioSynth(SAMPLE_RATE)
        total_samples = int(seconds) * SAMPLE_RATE
        per_frame, remainder = total_samples // total, total_samples % total
        log(f"Audio enabled: {SAMPLE_RATE} Hz, {per_frame} samples/frame (+{remainder} remainder)")
    proc, sink, path = None, None, mp4
    try:
        if use_ffmpeg:
            proc = subprocess.Popen(ffmpeg_args(size, fps, mp4), stdin=subprocess.PIPE)
            sink = proc.stdin
        else:
            path = os.path.join(output, f"{name}.rgb")
            sink = open(path, "wb")
        for step in range(total):
            eng.iterate_and_render()
            if step % int(fps) == 0:
                log(f"Frame {step + 1}/{total} ({step // int(fps)}s)...")
            sink.write(eng.download_render().tobytes())
            if synth is not None:
                chunks.append(synth.generate_frame(eng.render_features(), per_frame + (1 if step < remainder else 0)))
    finally:
        if sink is not None:
            sink.close()
        eng.close()
        if synth is not None:
            synth.close()
    if proc is not None:
        if proc.wait() != 0:
            raise RuntimeError(f"ffmpeg encoding failed (exit code {proc.returncode})")
        log(f"Video saved to: {mp4}")
    else:
        log(f"Raw rgb24 frames saved to: {path} (no ffmpeg on PATH; encode with: "
            f"{' '.join(ffmpeg_args(size, fps, mp4)).replace(' -i - ', f' -i {path} ')})")
    if with_sound and chunks:
        import numpy as np
        from .audio import SAMPLE_RATE, write_wav
        wav = os.path.join(output, f"{name}.wav")
        samples = np.concatenate(chunks)
        write_wav(wav, samples, SAMPLE_RATE, 2)
        log(f"Audio: {samples.size / (SAMPLE_RATE * 2.0):.1f}s of stereo PCM in {wav}")
        if proc is not None:
            muxed = os.path.join(output, f"{name}_muxed.mp4")
            if subprocess.run(mux_args(mp4, wav, muxed)).returncode == 0:
                os.replace(muxed, mp4)
                os.remove(wav)
                log(f"Audio muxed into: {mp4}")
            else:
                log(f"ffmpeg muxing failed. WAV kept at: {wav}")
    return path


def _main(argv=None):
    ap = argparse.ArgumentParser(prog="video", description="MaceLenia GPU simulation videos (ml-rs main.rs)")
    ap.add_argument("--creature")
    ap.add_argument("--seconds", type=int, default=60)
    ap.add_argument("--fps", type=int, default=60)
    ap.add_argument("--output", default="videos")
    ap.add_argument("--temp", type=float, default=1.0)
    ap.add_argument("--with-sound", action="store_true")
    args = ap.parse_args(argv)
    if args.seconds <= 0 or args.fps <= 0:
        print("--seconds and --fps must be positive", file=sys.stderr)
        return 2
    names = [args.creature] if args.creature else discover_creatures()
    if not names:
        print("No creatures found in seed/. Use --creature or run generate_kernel_json.py first.", file=sys.stderr)
        return 1
    if not args.creature:
        print(f"Found {len(names)} creatures in seed/. Generating videos for all...")
    for name in names:
        print(f"\n  Creature: {name}")
        render_video(name, args.seconds, args.fps, args.output, args.temp, with_sound=args.with_sound)
    return 0


if __name__ == "__main__":
    # Run as a file: load the package under its import name, then use its copy of this module.
    import importlib
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import __graft_entry__
    __graft_entry__.load_package()
    raise SystemExit(importlib.import_module("cella_nova_amd.video")._main())
