"""Frame harness: the reference's PNG frame dump and comparison (ml-rs/save_pngs.rs:19-168,
train/save_frames_png.py:202-211, train/compare_pngs.py:30-80), plus full-precision state dumps
so that two runs can be compared at f32 resolution instead of 8-bit gray levels.

  gray_frame(state)       save_pngs.rs:19-61: per-pixel f32 sum over channels (channel order),
                          min/max normalisation, x255, truncation to u8
  write_png / read_png    8-bit grayscale or RGB PNG (stdlib zlib; all five row filters decoded)
  save_frames(...)        frame 0 = the uploaded state, then one frame per step; each frame as
                          <prefix><step:04d>.png and the f32 state as <prefix><step:04d>.npy;
                          rgb=True also writes the render epilogue's frame (<prefix>rgb_<step>.png)
  compare_frames(a, b)    compare_pngs.py:30-80 per step: max gray difference, differing pixels,
                          "OK" when the max difference is <= 1 gray level; with both .npy dumps
                          present also the relative L-inf state difference

Command line (save_pngs.rs:66-168 layout): python cella-nova_amd/frames.py --creature NAME
reads seed/NAME.json and kernels/NAME_<size>.bin (size = seed_size rounded up to a power of two)
and writes pngs/cn_frame_XXXX.png for 50 steps; --compare PREFIX_A PREFIX_B compares two dumps.
"""
import argparse
import os
import struct
import zlib

import numpy as np

_PNG_SIG = b"\x89PNG\r\n\x1a\n"


def gray_frame(state):
    """(C, N, N) f32 state -> (N, N) uint8 frame as save_pngs.rs:19-61 computes it."""
    st = np.asarray(state, dtype=np.float32)
    s = np.zeros(st.shape[1:], dtype=np.float32)
    for c in range(st.shape[0]):  # f32 accumulation in channel order
        s = (s + st[c]).astype(np.float32)
    lo, hi = s.min(), s.max()
    rng = np.float32(hi - lo)
    if rng > 0:
        norm = ((s - lo) / rng).astype(np.float32)
    else:
        norm = np.zeros_like(s)
    return np.clip(np.trunc(norm * np.float32(255.0)), 0, 255).astype(np.uint8)


def _chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def write_png(path, img):
    """(H, W) uint8 grayscale or (H, W, 3) uint8 RGB -> PNG (filter 0 rows)."""
    a = np.ascontiguousarray(np.asarray(img, dtype=np.uint8))
    if a.ndim == 2:
        ctype, ch = 0, 1
    elif a.ndim == 3 and a.shape[2] == 3:
        ctype, ch = 2, 3
    else:
        raise ValueError("expected (H, W) or (H, W, 3) uint8")
    h, w = a.shape[:2]
    rows = a.reshape(h, w * ch)
    raw = b"".join(b"\x00" + rows[y].tobytes() for y in range(h))
    data = (_PNG_SIG + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
            + _chunk(b"IDAT", zlib.compress(raw, 6)) + _chunk(b"IEND", b""))
    with open(path, "wb") as f:
        f.write(data)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def read_png(path):
    """8-bit non-interlaced grayscale / RGB / RGBA PNG -> (H, W) or (H, W, C) uint8."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != _PNG_SIG:
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, hdr = 8, [], None
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if tag == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif tag == b"IDAT":
            idat.append(body)
        elif tag == b"IEND":
            break
    if hdr is None:
        raise ValueError(f"{path}: no IHDR")
    w, h, depth, ctype, _, _, interlace = hdr
    ch = {0: 1, 2: 3, 4: 2, 6: 4}.get(ctype)
    if depth != 8 or ch is None or interlace != 0:
        raise ValueError(f"{path}: only 8-bit non-interlaced gray/RGB/RGBA PNGs are supported")
    raw = zlib.decompress(b"".join(idat))
    stride = w * ch
    out = np.zeros((h, stride), dtype=np.uint8)
    prev = np.zeros(stride, dtype=np.int32)
    for y in range(h):
        ft = raw[y * (stride + 1)]
        line = np.frombuffer(raw, dtype=np.uint8, count=stride, offset=y * (stride + 1) + 1).astype(np.int32)
        cur = np.zeros(stride, dtype=np.int32)
        if ft == 0:
            cur = line
        elif ft == 2:
            cur = (line + prev) & 0xFF
        else:
            for i in range(stride):
                left = cur[i - ch] if i >= ch else 0
                if ft == 1:
                    cur[i] = (line[i] + left) & 0xFF
                elif ft == 3:
                    cur[i] = (line[i] + ((left + prev[i]) >> 1)) & 0xFF
                elif ft == 4:
                    ul = prev[i - ch] if i >= ch else 0
                    cur[i] = (line[i] + _paeth(left, prev[i], ul)) & 0xFF
                else:
                    raise ValueError(f"{path}: bad filter type {ft}")
        out[y] = cur
        prev = cur
    return out.reshape(h, w) if ch == 1 else out.reshape(h, w, ch)


def save_frames(engine, steps, outdir, prefix="cn_frame_", rgb=False, states=True):
    """Frame 0 (the current state) and one frame per step of a GpuMaceLenia engine."""
    os.makedirs(outdir, exist_ok=True)
    C, n = engine.num_channels, engine.n

    def dump(step):
        st = engine.download_all_channels().reshape(C, n, n)
        base = os.path.join(outdir, f"{prefix}{step:04d}")
        write_png(base + ".png", gray_frame(st))
        if states:
            np.save(base + ".npy", st)
        if rgb and step > 0:
            write_png(os.path.join(outdir, f"{prefix}rgb_{step:04d}.png"),
                      engine.download_render().reshape(n, n, 3))

    dump(0)
    for step in range(1, steps + 1):
        if rgb:
            engine.iterate_and_render()
        else:
            engine.iterate()
        dump(step)


def compare_frames(prefix_a, prefix_b, count=None):
    """Per-step comparison of two frame dumps (compare_pngs.py:30-80), plus the relative L-inf
    state difference where both runs saved .npy states.  Returns a list of dicts."""
    rows, step = [], 0
    while count is None or step < count:
        pa, pb = f"{prefix_a}{step:04d}.png", f"{prefix_b}{step:04d}.png"
        if not (os.path.exists(pa) and os.path.exists(pb)):
            break
        a, b = read_png(pa).astype(np.int64), read_png(pb).astype(np.int64)
        if a.shape != b.shape:
            raise ValueError(f"step {step}: frame shapes differ {a.shape} vs {b.shape}")
        d = np.abs(a - b)
        row = {"step": step, "sum_a": int(a.sum()), "sum_b": int(b.sum()), "max_diff": int(d.max()),
               "pixels_off": int((d > 0).sum()), "ok": bool(d.max() <= 1)}
        na, nb = f"{prefix_a}{step:04d}.npy", f"{prefix_b}{step:04d}.npy"
        if os.path.exists(na) and os.path.exists(nb):
            xa, xb = np.load(na).astype(np.float64), np.load(nb).astype(np.float64)
            row["rel_linf"] = float(np.abs(xa - xb).max() / max(np.abs(xb).max(), 1e-300))
        rows.append(row)
        step += 1
    return rows


def _main(argv=None):
    from .config import load_config, load_kernels
    from .engine import GpuMaceLenia

    ap = argparse.ArgumentParser(description="Save or compare DiffusionLenia frame dumps")
    ap.add_argument("--creature", help="creature name: seed/NAME.json and kernels/NAME_<size>.bin")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--out", default="pngs")
    ap.add_argument("--rgb", action="store_true", help="also save the render epilogue's RGB frames")
    ap.add_argument("--compare", nargs=2, metavar=("PREFIX_A", "PREFIX_B"))
    args = ap.parse_args(argv)
    if args.compare:
        rows = compare_frames(*args.compare)
        print(f"{'Step':>5}  {'A sum':>10}  {'B sum':>10}  {'Max diff':>9}  {'Px off':>7}  {'rel L-inf':>10}  Status")
        for r in rows:
            print(f"{r['step']:5d}  {r['sum_a']:10d}  {r['sum_b']:10d}  {r['max_diff']:9d}  {r['pixels_off']:7d}  "
                  f"{r.get('rel_linf', float('nan')):10.2e}  {'OK' if r['ok'] else 'DIFF'}")
        return 0 if rows and all(r["ok"] for r in rows) else 1
    if not args.creature:
        ap.error("--creature or --compare is required")
    cfg = load_config(os.path.join("seed", f"{args.creature}.json"))
    size = 1 << max(0, (int(cfg.seed_size) - 1).bit_length())
    kers = load_kernels(os.path.join("kernels", f"{args.creature}_{size}.bin"), cfg.num_kernels, size)
    eng = GpuMaceLenia(size, cfg.num_channels, cfg.num_kernels, cfg.c0, cfg.c1, cfg.growth_mu, cfg.growth_sigma,
                       cfg.growth_weights, temp=1.0)
    for k in range(cfg.num_kernels):
        eng.set_kernel(kers[k], k)
    for c in range(cfg.num_channels):
        eng.upload_channel(cfg.seed_channels[c], c)
    save_frames(eng, args.steps, args.out, rgb=args.rgb)
    eng.close()
    print(f"saved {args.steps + 1} frames to {args.out}/")
    return 0


if __name__ == "__main__":
    # Run as a file: load the package under its import name, then use its copy of this module.
    import importlib
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import __graft_entry__
    __graft_entry__.load_package()
    raise SystemExit(importlib.import_module("cella_nova_amd.frames")._main())
