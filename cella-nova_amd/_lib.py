"""ctypes binding of libcellanova.so (include/cellanova.h).

The shared library is built in-tree by ``make -C cella-nova_amd`` (or ``__graft_entry__.build()``).
Loading fails loudly when it is missing: there is no fallback implementation.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcellanova.so")
# An alternative build of the library can be loaded by path (e.g. a profiling build).
LIB_PATH = os.environ.get("CELLANOVA_LIB", LIB_PATH)

CN_OK = 0
STATUS_NAMES = {
    0: "CN_OK", 1: "CN_ERR_INVALID_ARG", 2: "CN_ERR_UNSUPPORTED", 3: "CN_ERR_HIP", 4: "CN_ERR_IO",
    5: "CN_ERR_PARSE", 6: "CN_ERR_OUT_OF_MEMORY", 7: "CN_ERR_STATE",
}


class CellaNovaError(RuntimeError):
    """A cn_* call returned a non-zero cn_status."""

    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class EngineDesc(ctypes.Structure):
    _fields_ = [
        ("n", c_uint32), ("num_channels", c_uint32), ("num_kernels", c_uint32),
        ("src", POINTER(c_uint32)), ("dst", POINTER(c_uint32)),
        ("mu", POINTER(c_float)), ("sigma", POINTER(c_float)), ("weights", POINTER(c_float)),
        ("temp", c_float), ("device", c_int),
    ]


class EngineInfo(ctypes.Structure):
    _fields_ = [("n", c_uint32), ("num_channels", c_uint32), ("num_kernels", c_uint32),
                ("steps_done", c_uint64), ("device_bytes", c_uint64), ("compact_kernels", c_uint32),
                ("segments", c_uint32), ("segment_radius", c_uint32), ("segment_length", c_uint32)]


class Config(ctypes.Structure):
    _fields_ = [
        ("seed_size", c_uint32), ("num_channels", c_uint32), ("num_kernels", c_uint32),
        ("seed_channels", POINTER(c_double)), ("seed_len", c_size_t),
        ("c0", POINTER(c_uint32)), ("c1", POINTER(c_uint32)),
        ("growth_mu", POINTER(c_float)), ("growth_sigma", POINTER(c_float)),
        ("growth_weights", POINTER(c_float)),
    ]


class CreatureParams(ctypes.Structure):
    _fields_ = [
        ("n", c_uint32), ("num_channels", c_uint32), ("num_kernels", c_uint32),
        ("src", POINTER(c_uint32)), ("dst", POINTER(c_uint32)), ("radius", POINTER(c_uint32)),
        ("patch", c_uint32), ("seed", c_uint64), ("creature", c_uint64),
    ]


class AdamParams(ctypes.Structure):
    _fields_ = [("lr", c_float), ("beta1", c_float), ("beta2", c_float), ("eps", c_float),
                ("mu_min", c_float), ("mu_max", c_float), ("sigma_min", c_float), ("weight_min", c_float)]


class Plan(ctypes.Structure):
    _fields_ = [("n", c_uint32), ("fwd_e", c_uint32), ("inv_e", c_uint32), ("col_e", c_uint32),
                ("col_tc", c_uint32), ("diff_th", c_uint32), ("col_split", c_uint32), ("col_full", c_uint32)]


# name -> (restype, argtypes); every symbol declared in include/cellanova.h
SIGNATURES = {
    "cn_engine_create": (c_int, [POINTER(EngineDesc), POINTER(c_void_p)]),
    "cn_engine_destroy": (c_int, [c_void_p]),
    "cn_set_kernel": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    "cn_upload_channel_f64": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    "cn_upload_channel_f32": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    "cn_iterate": (c_int, [c_void_p]),
    "cn_run_steps": (c_int, [c_void_p, c_uint64]),
    "cn_sync": (c_int, [c_void_p]),
    "cn_iterate_on_stream": (c_int, [c_void_p, c_void_p]),
    "cn_engine_stream": (c_int, [c_void_p, POINTER(c_void_p)]),
    "cn_creature_state_ptr": (c_int, [c_void_p, c_uint32, POINTER(c_void_p), POINTER(c_size_t)]),
    "cn_timed_iterate": (c_int, [c_void_p, POINTER(c_double)]),
    "cn_profile_kernels": (c_int, [c_void_p, c_uint64, c_void_p, c_size_t]),
    "cn_download_channels": (c_int, [c_void_p, c_void_p, c_size_t]),
    "cn_download_conv_real": (c_int, [c_void_p, c_void_p, c_size_t]),
    "cn_mass_fp64": (c_int, [c_void_p, c_void_p, c_size_t]),
    "cn_mass_fp64_noalloc": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_void_p]),
    "cn_iterate_and_render": (c_int, [c_void_p]),
    "cn_download_render": (c_int, [c_void_p, c_void_p, c_size_t]),
    "cn_creature_download_render": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    "cn_engine_get_info": (c_int, [c_void_p, POINTER(EngineInfo)]),
    "cn_last_error": (c_char_p, []),
    "cn_load_config": (c_int, [c_char_p, POINTER(POINTER(Config))]),
    "cn_free_config": (None, [POINTER(Config)]),
    "cn_load_kernels": (c_int, [c_char_p, c_uint32, c_uint32, c_void_p, c_size_t]),
    "cn_write_kernels": (c_int, [c_char_p, c_void_p, c_uint32, c_uint32]),
    "cn_write_config": (c_int, [c_char_p, POINTER(Config)]),
    "cn_creature_generate": (c_int, [c_void_p, c_uint32, POINTER(CreatureParams)]),
    "cn_generate_example_creature": (c_int, [c_uint32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                             c_void_p, c_void_p]),
    "cn_creature_download_kernels": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    "cn_generate_creature": (c_int, [POINTER(CreatureParams), c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    "cn_debug_fail_next_hip": (None, [c_int]),
    "cn_debug_fail_hip_at": (None, [c_int]),
    # creature search on the device
    "cn_batch_get_growth": (c_int, [c_void_p, c_void_p, c_size_t]),
    "cn_batch_set_growth": (c_int, [c_void_p, c_void_p, c_size_t]),
    "cn_batch_get_param_grads": (c_int, [c_void_p, c_void_p, c_size_t]),
    "cn_adam_step": (c_int, [c_void_p, POINTER(AdamParams)]),
    "cn_adam_reset": (c_int, [c_void_p]),
    # plans / autotuner
    "cn_autotune": (c_int, [c_void_p, c_uint32]),
    "cn_engine_size_classes": (c_uint32, [c_void_p]),
    "cn_engine_get_plan": (c_int, [c_void_p, c_uint32, POINTER(Plan)]),
    "cn_engine_set_plan": (c_int, [c_void_p, POINTER(Plan)]),
    # batches
    "cn_batch_create": (c_int, [POINTER(EngineDesc), c_uint32, c_int, POINTER(c_void_p)]),
    "cn_creature_set_kernel": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_size_t]),
    "cn_creature_upload_channel_f32": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_size_t]),
    "cn_creature_download_channels": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    "cn_creature_mass_fp64": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    "cn_creature_set_growth": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p, c_void_p, c_size_t]),
    "cn_creature_state_from_device": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    "cn_creature_state_to_device": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    # rollouts
    "cn_creature_set_target": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    "cn_rollout_grad": (c_int, [c_void_p, c_uint32, c_uint32]),
    "cn_rollout_forward": (c_int, [c_void_p, c_uint32, c_uint32]),
    "cn_creature_rollout_final": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t, c_int]),
    "cn_rollout_backward": (c_int, [c_void_p, POINTER(c_void_p), c_uint32]),
    "cn_creature_get_loss": (c_int, [c_void_p, c_uint32, POINTER(c_double)]),
    "cn_creature_get_grads": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p,
                                      c_size_t, c_int]),
    # audio consumer of the rendered frames
    "cn_creature_render_features": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p]),
    "cn_render_features": (c_int, [c_void_p, c_void_p, c_void_p]),
    "cn_audio_extract_features": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p]),
    "cn_audio_create": (c_int, [c_uint32, POINTER(c_void_p)]),
    "cn_audio_destroy": (c_int, [c_void_p]),
    "cn_audio_generate_frame": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_size_t]),
    "cn_audio_activity": (c_int, [c_void_p, POINTER(c_float)]),
    "cn_write_wav": (c_int, [c_char_p, c_void_p, c_size_t, c_uint32, ctypes.c_uint16]),
}

_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C cella-nova_amd` or __graft_entry__.build()")
        # One HIP runtime per process: PyTorch's ROCm wheel carries its own libamdhip64 (same
        # soname).  Loading torch first lets this library bind to that copy, so device pointers
        # and streams are shared with torch; loaded the other way round, torch would load a
        # second runtime and fail to find the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def check(status):
    """Raise CellaNovaError for a non-zero status."""
    if status != CN_OK:
        msg = lib().cn_last_error()
        raise CellaNovaError(status, msg.decode() if msg else "")
    return status
