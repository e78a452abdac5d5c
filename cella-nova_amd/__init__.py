"""MI355X-native DiffusionLenia step engine: the drop-in for the reference's `GpuMaceLenia`
orchestrator and its loader, over the C ABI in include/cellanova.h (libcellanova.so)."""
from ._lib import CellaNovaError, LIB_PATH, lib
from .config import SimConfig, load_config, load_kernels, write_config, write_kernels
from .creatures import (baseline_creature, channel_maps, creature_spec, example_creature, generate_creature,
                        make_engine)
from .engine import CreatureBatch, GpuMaceLenia
from . import frames, video

__all__ = ["CellaNovaError", "CreatureBatch", "frames", "video", "GpuMaceLenia", "LIB_PATH", "SimConfig", "baseline_creature", "channel_maps",
           "creature_spec", "example_creature", "generate_creature", "lib", "load_config", "load_kernels", "make_engine",
           "write_config", "write_kernels"]
