"""Frame harness (save_pngs.rs:19-61, compare_pngs.py:30-80): PNG codec, gray-frame formula and
the comparison report, on the CPU."""
import struct
import zlib

import numpy as np


def _png_with_filters(img, filters):
    """Independent encoder for the PNG row filters (PNG spec section 6): Sub, Up, Average, Paeth."""
    h, w = img.shape
    raw = b""
    prev = np.zeros(w, np.int32)
    for y in range(h):
        x = img[y].astype(np.int32)
        ft = filters[y % len(filters)]
        out = np.zeros(w, np.int32)
        for i in range(w):
            a = x[i - 1] if i else 0
            b = prev[i]
            c = prev[i - 1] if i else 0
            if ft == 0:
                pred = 0
            elif ft == 1:
                pred = a
            elif ft == 2:
                pred = b
            elif ft == 3:
                pred = (a + b) // 2
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            out[i] = (x[i] - pred) & 0xFF
        raw += bytes([ft]) + out.astype(np.uint8).tobytes()
        prev = x

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 0, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def test_png_round_trip_gray_and_rgb(cn, tmp_path):
    rng = np.random.default_rng(0)
    g = rng.integers(0, 256, (17, 23), dtype=np.uint8)
    c = rng.integers(0, 256, (9, 5, 3), dtype=np.uint8)
    cn.frames.write_png(tmp_path / "g.png", g)
    cn.frames.write_png(tmp_path / "c.png", c)
    assert np.array_equal(cn.frames.read_png(tmp_path / "g.png"), g)
    assert np.array_equal(cn.frames.read_png(tmp_path / "c.png"), c)
    assert (tmp_path / "g.png").read_bytes()[:8] == b"\x89PNG\r\n\x1a\n"


def test_png_reader_decodes_every_row_filter(cn, tmp_path):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (10, 13), dtype=np.uint8)
    p = tmp_path / "f.png"
    p.write_bytes(_png_with_filters(img, [0, 1, 2, 3, 4]))
    assert np.array_equal(cn.frames.read_png(p), img)


def test_gray_frame_formula(cn):
    f = cn.frames.gray_frame
    assert not f(np.full((3, 4, 4), 0.3, np.float32)).any()          # zero range -> black
    st = np.zeros((2, 2, 2), np.float32)
    st[0, 0, 1] = 1.0                                                  # channel sums 0, 1, 2
    st[:, 1, 1] = 1.0
    out = f(st)
    assert out.tolist() == [[0, 127], [0, 255]]                        # 0.5 * 255 = 127.5 truncates
    assert out.dtype == np.uint8


def test_compare_frames_report(cn, tmp_path):
    a = np.zeros((8, 8), np.uint8)
    b = a.copy()
    b[2, 3] = 1
    c = a.copy()
    c[0, 0] = 5
    for step, (x, y) in enumerate([(a, a), (a, b), (a, c)]):
        cn.frames.write_png(tmp_path / f"A_{step:04d}.png", x)
        cn.frames.write_png(tmp_path / f"B_{step:04d}.png", y)
    np.save(tmp_path / "A_0000.npy", np.ones((1, 8, 8), np.float32))
    np.save(tmp_path / "B_0000.npy", np.full((1, 8, 8), 1.5, np.float32))
    rows = cn.frames.compare_frames(str(tmp_path / "A_"), str(tmp_path / "B_"))
    assert [r["max_diff"] for r in rows] == [0, 1, 5]
    assert [r["ok"] for r in rows] == [True, True, False]
    assert rows[1]["pixels_off"] == 1 and abs(rows[0]["rel_linf"] - 1 / 3) < 1e-12
