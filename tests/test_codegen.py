"""Machine-code properties of the step kernels for gfx950 (compiled here, no GPU needed).

Every device-memory access of the hot kernels must be a global (address space 1) load or store:
FLAT instructions also count on the LDS/scalar-memory counter, so in a kernel that waits on LDS
exchanges each such wait would also drain all device-memory loads and stores in flight.  The
kernels' pointers come from the creature descriptor, which the compiler cannot prove to be
global memory unless the source says so (types.h: gptr / to_global).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cella-nova_amd")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HOT = ("k_row_fwd", "k_col_fused", "k_col_segments", "k_row_inv_growth", "k_diffusion", "k_diffusion_adj",
       "k_growth_adj_row", "k_col_adj", "k_row_inv_add", "k_gen_rows", "k_gen_cols")


_ASM = {}


def _device_asm(src, tmp_path):
    """gfx950 assembly of one source file (compiled once per test session)."""
    if src in _ASM:
        return _ASM[src]
    if not (os.path.exists(HIPCC) or shutil.which(HIPCC)):
        pytest.skip("hipcc not available")
    out = tmp_path / (os.path.basename(src) + ".s")
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", f"-I{os.path.join(ROOT, 'include')}", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S",
           os.path.join(PKG, "csrc", src), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    _ASM[src] = out.read_text()
    return _ASM[src]


def _kernels(asm):
    """{mangled name: body} for the cn:: kernels of an assembly file."""
    out = {}
    for m in re.finditer(r"^(_ZN2cn\w+):", asm, re.M):
        end = asm.find(".Lfunc_end", m.end())
        out[m.group(1)] = asm[m.end():end]
    return out


@pytest.mark.parametrize("src", ["step_kernels.hip", "adjoint_kernels.hip", "gen_kernels.hip"])
def test_hot_kernels_use_global_memory_instructions(src, tmp_path):
    kernels = _kernels(_device_asm(src, tmp_path))
    hot = {k: b for k, b in kernels.items() if any(h in k for h in HOT)}
    assert hot, "no kernels found in the assembly"
    flat = {k: len(re.findall(r"\bflat_(?:load|store)", b)) for k, b in hot.items()}
    bad = {k: n for k, n in flat.items() if n}
    assert not bad, f"FLAT memory instructions in {bad}"
    assert all(re.search(r"\bglobal_(?:load|store)", b) for b in hot.values())


@pytest.mark.parametrize("src", ["step_kernels.hip", "adjoint_kernels.hip"])
def test_benchmark_size_kernels_do_not_spill(src, tmp_path):
    """No instance for grid sizes 512..2048 (the batched and single-creature benchmark sizes)
    spills registers to scratch memory: a spill puts slow scratch loads and stores in the
    transforms' inner loops (a single one cost the 2048^2 column pass 15 % once)."""
    kernels = _kernels(_device_asm(src, tmp_path))
    sized = {k: b for k, b in kernels.items() if re.search(r"ILi(512|1024|2048)E", k)}
    assert sized, "no sized kernels found"
    spills = {k: len(re.findall(r"\bscratch_(?:load|store)", b)) for k, b in sized.items()}
    bad = {k: n for k, n in spills.items() if n}
    assert not bad, f"scratch spills in {bad}"
