"""Loader parity with ml-rs/config.rs on synthetic files with hand-computed expectations."""
import json
import struct

import numpy as np
import pytest


def _cfg(n=4, C=2, K=3, **over):
    d = {"seed_size": n, "num_channels": C, "num_kernels": K,
         "seed_channels": [[float(c * 100 + i) for i in range(n * n)] for c in range(C)],
         "c0": [0, 1, 1], "c1": [1, 0, 1], "growth_mu": [0.1, 0.2, 0.3], "growth_sigma": [0.01, 0.02, 0.03],
         "growth_weights": [0.5, 0.25, 1], "global_r": 10.0, "radii": [0.25, 0.3, 0.4]}
    d.update(over)
    return d


def test_load_config_swaps_c0_c1(cn, tmp_path):
    p = tmp_path / "c.json"
    p.write_text(json.dumps(_cfg()))
    c = cn.load_config(p)
    assert (c.seed_size, c.num_channels, c.num_kernels) == (4, 2, 3)
    assert list(c.c0) == [1, 0, 1]  # JSON c1 (config.rs:40)
    assert list(c.c1) == [0, 1, 1]  # JSON c0
    assert np.allclose(c.growth_mu, [0.1, 0.2, 0.3]) and c.growth_mu.dtype == np.float32
    assert np.allclose(c.growth_weights, [0.5, 0.25, 1.0])
    assert c.seed_channels[1][5] == 105.0 and c.seed_channels[0].dtype == np.float64


def test_load_config_whitespace_unicode_and_order(cn, tmp_path):
    txt = json.dumps(_cfg(), indent=3, ensure_ascii=True)
    txt = txt.replace('"global_r"', '"gl\\u00f6bal_r"')
    p = tmp_path / "c.json"
    p.write_text(txt)
    c = cn.load_config(p)
    assert c.num_kernels == 3


@pytest.mark.parametrize("mutate,status", [
    (lambda d: d.pop("growth_mu"), 5),              # missing field
    (lambda d: d.update(c0=[0, 1.5, 1]), 5),         # non-integer u32
    (lambda d: d.update(c0=[0, -1, 1]), 5),          # negative u32
    (lambda d: d.update(growth_sigma=[0.1]), 5),     # wrong length
    (lambda d: d.update(seed_channels=[[0.0] * 15, [0.0] * 16]), 5),  # seed not seed_size^2
    (lambda d: d.update(num_channels=3), 5),         # channels vs seed count
])
def test_load_config_rejects(cn, tmp_path, mutate, status):
    d = _cfg()
    mutate(d)
    p = tmp_path / "c.json"
    p.write_text(json.dumps(d))
    with pytest.raises(cn.CellaNovaError) as ei:
        cn.load_config(p)
    assert ei.value.status == status


def test_load_config_malformed_and_missing(cn, tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"seed_size": 4, "num_channels": ')
    with pytest.raises(cn.CellaNovaError) as ei:
        cn.load_config(p)
    assert ei.value.status == 5
    # a number beyond the f64 range is a parse error (serde_json refuses it), not an infinity
    p.write_text(json.dumps(_cfg()).replace("105.0", "1e400"))
    with pytest.raises(cn.CellaNovaError) as ei:
        cn.load_config(p)
    assert ei.value.status == 5 and "out of range" in str(ei.value)
    p.write_text(json.dumps(_cfg()).replace("105.0", "1e-400"))  # underflow reads as 0, as serde_json
    assert cn.load_config(p).seed_channels[1][5] == 0.0
    with pytest.raises(cn.CellaNovaError) as ei:
        cn.load_config(tmp_path / "nope.json")
    assert ei.value.status == 4


def _write_bin(path, kernels):
    with open(path, "wb") as f:
        for k in kernels:
            f.write(np.ascontiguousarray(k.astype(np.complex64)).view(np.float32).astype("<f4").tobytes())


def test_load_kernels_same_size(cn, tmp_path):
    rng = np.random.default_rng(0)
    ks = [rng.standard_normal(64) + 1j * rng.standard_normal(64) for _ in range(3)]
    _write_bin(tmp_path / "k.bin", ks)
    out = cn.load_kernels(tmp_path / "k.bin", 3, 8)
    for i in range(3):
        assert np.array_equal(out[i], ks[i].astype(np.complex64))


def test_load_kernels_spectral_zero_padding(cn, tmp_path):
    """S=4 -> target 8: fftshift, centre pad by 2, shift by 4 (config.rs:90-116).  A constant
    spectrum becomes ones at the frequencies |f| <= 2 that the 4-point grid holds
    (index sets {0,1,6,7} in each axis for -2..1 after the shifts)."""
    S, T = 4, 8
    _write_bin(tmp_path / "k.bin", [np.ones(S * S)])
    out = cn.load_kernels(tmp_path / "k.bin", 1, T).reshape(T, T)
    idx = [0, 1, 6, 7]
    want = np.zeros((T, T), np.complex64)
    want[np.ix_(idx, idx)] = 1
    assert np.array_equal(out, want)
    # frequency-preserving: a spectrum with a marker at (1, 3) (i.e. f = (1, -1)) lands at (1, 7)
    m = np.zeros((S, S), complex)
    m[1, 3] = 2 + 1j
    _write_bin(tmp_path / "m.bin", [m.reshape(-1)])
    out = cn.load_kernels(tmp_path / "m.bin", 1, T).reshape(T, T)
    assert out[1, 7] == np.complex64(2 + 1j) and np.count_nonzero(out) == 1


def test_load_kernels_rejects_bad_files(cn, tmp_path):
    (tmp_path / "odd.bin").write_bytes(b"\0" * (3 * 8 * 16 + 4))
    with pytest.raises(cn.CellaNovaError) as ei:
        cn.load_kernels(tmp_path / "odd.bin", 3, 16)
    assert ei.value.status == 4
    (tmp_path / "ns.bin").write_bytes(b"\0" * (2 * 8 * 15))  # 15 values per kernel: not a square
    with pytest.raises(cn.CellaNovaError) as ei:
        cn.load_kernels(tmp_path / "ns.bin", 2, 16)
    assert ei.value.status == 4
    _write_bin(tmp_path / "big.bin", [np.ones(64)])
    with pytest.raises(cn.CellaNovaError) as ei:
        cn.load_kernels(tmp_path / "big.bin", 1, 4)  # target < stored size
    assert ei.value.status == 1
    with pytest.raises(cn.CellaNovaError) as ei:
        cn.load_kernels(tmp_path / "missing.bin", 1, 4)
    assert ei.value.status == 4


# --- writers (train/generate_kernel_json.py:171-213) ------------------------------------------
def test_write_kernels_is_the_reference_bin_format(cn, tmp_path):
    rng = np.random.default_rng(3)
    k = (rng.standard_normal((3, 64)) + 1j * rng.standard_normal((3, 64))).astype(np.complex64)
    p = tmp_path / "k.bin"
    cn.write_kernels(p, k)
    raw = p.read_bytes()
    assert len(raw) == 3 * 64 * 8
    # struct.pack("ff", re, im) per value, kernel-major then row-major
    want = b"".join(struct.pack("<ff", float(v.real), float(v.imag)) for v in k.reshape(-1))
    assert raw == want
    assert np.array_equal(cn.load_kernels(p, 3, 8), k)
    padded = cn.load_kernels(p, 3, 16)            # and the loader's spectral zero-pad path reads it
    assert padded.shape == (3, 256)


def test_write_config_round_trip_and_json_convention(cn, tmp_path):
    rng = np.random.default_rng(4)
    C, K, n = 3, 5, 4
    cfg = cn.SimConfig(seed_size=n, num_channels=C, num_kernels=K,
                       seed_channels=[rng.random(n * n) for _ in range(C)],
                       c0=np.array([0, 1, 2, 0, 1], np.uint32), c1=np.array([2, 2, 0, 1, 1], np.uint32),
                       growth_mu=rng.random(K).astype(np.float32), growth_sigma=rng.random(K).astype(np.float32),
                       growth_weights=rng.random(K).astype(np.float32))
    p = tmp_path / "c.json"
    cn.write_config(p, cfg)
    doc = json.loads(p.read_text())
    assert sorted(doc) == sorted(["seed_size", "num_channels", "num_kernels", "seed_channels", "c0", "c1",
                                  "growth_mu", "growth_sigma", "growth_weights"])
    assert doc["c0"] == cfg.c1.tolist() and doc["c1"] == cfg.c0.tolist()   # JSON convention (config.rs:40)
    assert np.array_equal(np.array(doc["seed_channels"]), np.array(cfg.seed_channels))  # 17 digits: exact
    back = cn.load_config(p)
    assert back.seed_size == n and back.num_channels == C and back.num_kernels == K
    assert np.array_equal(back.c0, cfg.c0) and np.array_equal(back.c1, cfg.c1)
    for a, b in ((back.growth_mu, cfg.growth_mu), (back.growth_sigma, cfg.growth_sigma),
                 (back.growth_weights, cfg.growth_weights)):
        assert np.array_equal(a, b)
    assert all(np.array_equal(a, b) for a, b in zip(back.seed_channels, cfg.seed_channels))


def test_writers_reject_bad_arguments(cn, tmp_path):
    with pytest.raises(cn.CellaNovaError):
        cn.write_kernels(tmp_path / "no_dir" / "k.bin", np.zeros((1, 16), np.complex64))
    cfg = cn.SimConfig(seed_size=4, num_channels=1, num_kernels=1, seed_channels=[np.full(16, np.nan)],
                       c0=np.zeros(1, np.uint32), c1=np.zeros(1, np.uint32), growth_mu=np.ones(1, np.float32),
                       growth_sigma=np.ones(1, np.float32), growth_weights=np.ones(1, np.float32))
    with pytest.raises(cn.CellaNovaError, match="not finite"):
        cn.write_config(tmp_path / "c.json", cfg)


_OOM_CHILD = r"""
import ctypes, resource, sys
lib = ctypes.CDLL(sys.argv[1])
lib.cn_last_error.restype = ctypes.c_char_p
vm = [int(l.split()[1]) for l in open("/proc/self/status") if l.startswith("VmSize:")][0] * 1024
resource.setrlimit(resource.RLIMIT_AS, (vm + (512 << 20), resource.getrlimit(resource.RLIMIT_AS)[1]))
out = ctypes.c_void_p()
st = lib.cn_load_config(sys.argv[2].encode(), ctypes.byref(out))
print(st, lib.cn_last_error().decode(), out.value)
"""


def test_load_config_host_allocation_failure_is_a_status(cn, tmp_path):
    """A config larger than the host can hold: cn_load_config returns CN_ERR_OUT_OF_MEMORY
    instead of letting std::bad_alloc cross the C ABI.  A sparse 1 TiB file read in a child
    process whose address space is capped 512 MiB above its size at the call."""
    import subprocess
    import sys
    big = tmp_path / "huge.json"
    try:
        with open(big, "wb") as f:
            f.truncate(1 << 40)
    except OSError:
        pytest.skip("no sparse files on this filesystem")
    r = subprocess.run([sys.executable, "-c", _OOM_CHILD, cn.LIB_PATH, str(big)], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    assert r.stdout.decode().split() == ["6", "host", "allocation", "failed", "None"]  # CN_ERR_OUT_OF_MEMORY
