"""C ABI checks that need no GPU: every declared symbol is exported, host-side validation
rejects bad arguments before any HIP call, injected HIP failures surface as CN_ERR_HIP, and
the synthetic creature generator follows its stated distributions."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cellanova.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:cn_status|void|const char\*|uint32_t)\s+(cn_\w+)\s*\(", text, re.M)))


def test_header_symbols_are_exported_and_bound(cn):
    syms = declared_symbols()
    assert len(syms) >= 20
    h = ctypes.CDLL(cn.LIB_PATH)
    for s in syms:
        assert hasattr(h, s), s
    from cella_nova_amd._lib import SIGNATURES
    assert sorted(SIGNATURES) == syms


def _desc(cn, n=64, C=3, src=None, dst=None, mu=None, sigma=None, w=None, temp=1.0):
    from cella_nova_amd._lib import EngineDesc
    K = 9 if src is None else len(src)
    keep = [np.ascontiguousarray(x, dtype=dt) for x, dt in (
        (src if src is not None else [0, 1, 2] * 3, np.uint32),
        (dst if dst is not None else [0, 0, 0, 1, 1, 1, 2, 2, 2], np.uint32),
        (mu if mu is not None else [0.1] * K, np.float32),
        (sigma if sigma is not None else [0.04] * K, np.float32),
        (w if w is not None else [1 / 3] * K, np.float32))]
    P = ctypes.POINTER
    d = EngineDesc(n, C, K, keep[0].ctypes.data_as(P(ctypes.c_uint32)), keep[1].ctypes.data_as(P(ctypes.c_uint32)),
                   keep[2].ctypes.data_as(P(ctypes.c_float)), keep[3].ctypes.data_as(P(ctypes.c_float)),
                   keep[4].ctypes.data_as(P(ctypes.c_float)), temp, 0)
    return d, keep


@pytest.mark.parametrize("kw,status", [
    (dict(n=100), 2), (dict(n=4096), 2), (dict(n=8), 2), (dict(C=17), 2), (dict(C=0), 2),
    (dict(src=[0, 3, 1], dst=[0, 0, 0], C=3), 1), (dict(src=[0], dst=[5], C=3), 1),
    (dict(sigma=[0.0] * 9), 1), (dict(sigma=[-1.0] * 9), 1), (dict(mu=[np.nan] * 9), 1),
    (dict(w=[np.inf] * 9), 1), (dict(temp=float("nan")), 1),
])
def test_engine_create_validates_before_hip(cn, kw, status):
    lib = cn.lib()
    d, keep = _desc(cn, **kw)
    h = ctypes.c_void_p()
    assert lib.cn_engine_create(ctypes.byref(d), ctypes.byref(h)) == status
    assert h.value is None
    assert len(lib.cn_last_error()) > 0


def test_null_arguments_rejected(cn):
    lib = cn.lib()
    h = ctypes.c_void_p()
    assert lib.cn_engine_create(None, ctypes.byref(h)) == 1
    assert lib.cn_iterate(None) == 1
    assert lib.cn_run_steps(None, 3) == 1
    assert lib.cn_set_kernel(None, 0, None, 0) == 1
    assert lib.cn_download_channels(None, None, 0) == 1
    assert lib.cn_mass_fp64(None, None, 0) == 1
    assert lib.cn_engine_destroy(None) == 1
    assert lib.cn_load_kernels(None, 1, 16, None, 0) == 1


def test_batch_size_limit_checked_before_hip(cn):
    """More creatures than a launch grid's z dimension holds is refused up front (the
    descriptors beyond the first are not read)."""
    lib = cn.lib()
    d, keep = _desc(cn)
    h = ctypes.c_void_p()
    assert lib.cn_batch_create(ctypes.byref(d), 65536, 0, ctypes.byref(h)) == 2
    assert b"65535" in lib.cn_last_error()
    assert h.value is None


def test_injected_hip_failure_is_reported(cn):
    lib = cn.lib()
    d, keep = _desc(cn)
    h = ctypes.c_void_p()
    lib.cn_debug_fail_next_hip(1)
    assert lib.cn_engine_create(ctypes.byref(d), ctypes.byref(h)) == 3
    assert b"injected" in lib.cn_last_error()
    assert h.value is None
    lib.cn_debug_fail_next_hip(0)


def test_python_mirror_raises(cn):
    with pytest.raises(cn.CellaNovaError) as ei:
        cn.GpuMaceLenia(100, 3, 9, [0, 1, 2] * 3, [0] * 9, [0.1] * 9, [0.04] * 9, [0.3] * 9)
    assert ei.value.status == 2
    with pytest.raises(ValueError):
        cn.GpuMaceLenia((64, 64, 3), 3, 9, [0, 1, 2] * 3, [0] * 9, [0.1] * 9, [0.04] * 9, [0.3] * 9)


# ---------------------------------------------------------------- creature generator
def test_creature_distributions_and_determinism(cn):
    a = cn.baseline_creature(64, num_channels=3, radius=9, seed=3, patch=32, spatial=True)
    b = cn.baseline_creature(64, num_channels=3, radius=9, seed=3, patch=32, spatial=True)
    c = cn.baseline_creature(64, num_channels=3, radius=9, seed=4, patch=32)
    for key in ("kernels_fft", "state", "mu", "sigma", "weights"):
        assert np.array_equal(a[key], b[key])
    assert not np.array_equal(a["state"], c["state"])
    assert np.all((a["mu"] >= 0.05) & (a["mu"] < 0.15))
    assert np.all((a["sigma"] >= 0.03) & (a["sigma"] < 0.06))
    for dst in range(3):
        assert abs(a["weights"][a["dst"] == dst].sum() - 1.0) < 1e-6  # Dirichlet(1) per destination
    sp = a["spatial"]
    assert np.allclose(sp.sum(axis=(1, 2)), 1.0, atol=1e-12)  # normalised in FP64
    assert np.all(sp >= 0)
    y, x = np.mgrid[0:64, 0:64]
    r = np.hypot(np.minimum(y, 64 - y), np.minimum(x, 64 - x))
    assert np.all(sp[:, r >= 9] == 0)  # support inside the radius, origin at (0, 0)
    assert np.array_equal(sp, sp[:, np.r_[0, 63:0:-1], :])  # even in y
    spec = np.fft.fft2(sp).reshape(9, -1)
    assert np.abs(spec - a["kernels_fft"]).max() < 1e-6  # spectrum = FFT of the kernel, complex64
    st = a["state"]
    assert np.all(st[:, :16, :] == 0) and np.all(st[:, 48:, :] == 0)
    assert np.all((st[:, 16:48, 16:48] >= 0) & (st[:, 16:48, 16:48] < 1))


def test_lenia_shell_closed_form(cn):
    """One ring: K(r) = beta * exp(4 - 1/(r(1-r))) for r = dist/R in (0, 1) (before normalisation)."""
    a = cn.generate_creature(64, 1, 1, [0], [0], 20, patch=8, seed=0, spatial=True)
    sp = a["spatial"][0]
    y, x = np.mgrid[0:64, 0:64]
    r = np.hypot(np.minimum(y, 64 - y), np.minimum(x, 64 - x)) / 20.0
    nz = sp > 0
    # the ratio K / core is constant per ring; with B rings the core argument is f = B*r - floor(B*r)
    for B in (1, 2, 3):
        s = B * r
        f = s - np.floor(s)
        core = np.where((r < 1) & (f > 0), np.exp(4 - 1 / np.maximum(f * (1 - f), 1e-300)), 0.0)
        if np.array_equal(core > 0, nz):
            ring = np.floor(s)[nz].astype(int)
            ratio = sp[nz] / core[nz]
            for i in range(B):
                vals = ratio[ring == i]
                assert vals.max() - vals.min() < 1e-12 * max(vals.max(), 1e-300)
            return
    pytest.fail("spatial kernel does not match a 1-3 ring Lenia shell")


def test_creature_rejects_bad_params(cn):
    with pytest.raises(cn.CellaNovaError):
        cn.generate_creature(100, 1, 1, [0], [0], 5)
    with pytest.raises(cn.CellaNovaError):
        cn.generate_creature(64, 2, 1, [2], [0], 5)


def test_python_mirror_checks_tensor_arguments(cn):
    """Device-pointer entry points copy numel * 4 raw bytes: the Python mirror rejects anything
    but contiguous float32 device tensors of the right size before calling them (no GPU needed:
    the checks come first)."""
    torch = pytest.importorskip("torch")
    from cella_nova_amd.engine import _dev_f32, _host_f32
    with pytest.raises(TypeError):
        _dev_f32(torch.zeros(4), 4, "t")                      # host tensor
    with pytest.raises(TypeError):
        _dev_f32(np.zeros(4, np.float32), 4, "t")             # numpy array
    with pytest.raises(TypeError):
        _host_f32(np.zeros(4, np.float64), 4, "a")            # wrong dtype
    with pytest.raises(TypeError):
        _host_f32(np.zeros(8, np.float32)[::2], 4, "a")       # not contiguous
    with pytest.raises(TypeError):
        _host_f32(np.zeros(5, np.float32), 4, "a")            # wrong size
    assert _host_f32(np.zeros(4, np.float32), 4, "a") != 0


def test_stream_and_state_pointer_entries_validate(cn):
    lib = cn.lib()
    p = ctypes.c_void_p()
    assert lib.cn_engine_stream(None, ctypes.byref(p)) == 1
    assert lib.cn_iterate_on_stream(None, None) == 1
    n = ctypes.c_size_t()
    assert lib.cn_creature_state_ptr(None, 0, ctypes.byref(p), ctypes.byref(n)) == 1
    assert len(lib.cn_last_error()) > 0


def test_adam_and_batch_parameter_entries_validate(cn):
    from cella_nova_amd._lib import AdamParams
    lib = cn.lib()
    p = AdamParams(5e-3, 0.9, 0.999, 1e-8, 0.0, 2.0, 1e-3, 0.0)
    assert lib.cn_adam_step(None, ctypes.byref(p)) == 1
    assert lib.cn_adam_reset(None) == 1
    buf = np.zeros(27, np.float32)
    assert lib.cn_batch_get_growth(None, buf.ctypes.data, buf.size) == 1
    assert lib.cn_batch_set_growth(None, buf.ctypes.data, buf.size) == 1
    assert lib.cn_batch_get_param_grads(None, buf.ctypes.data, buf.size) == 1
    assert len(lib.cn_last_error()) > 0
