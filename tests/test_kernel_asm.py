"""Static checks of the compiled matrix-core tile kernel (CPU only: hipcc
cross-compiles gfx950 assembly without a GPU).

A matrix product (v_mfma_*) reads its accumulator over several passes and
writes its result passes later; the compiler inserts the wait states that
VALU instructions reading or writing those registers need, but it does not
check inline assembly against them. The tile kernel therefore keeps every
vector instruction compiler-visible: no inline-assembly vector instruction
may appear in it.
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "turbo-metrics_amd" / "csrc" / "ssim2_fast.hip"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
KERNEL = "_ZN5ssim216k_fast_tile_mfmaEPKNS_8PairDescEPKiS4_"


@pytest.fixture(scope="module")
def tile_asm(tmp_path_factory):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("asm") / "fast.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           f"-I{ROOT / 'include'}", "--cuda-device-only", "-S", "-o", str(out), str(SRC)]
    subprocess.run(cmd, check=True, capture_output=True)
    text = out.read_text().splitlines()
    start = next(i for i, ln in enumerate(text) if ln.startswith(KERNEL + ":"))
    end = next(i for i in range(start, len(text)) if text[i].startswith(".Lfunc_end"))
    return text[start:end]


def test_tile_kernel_uses_matrix_cores(tile_asm):
    assert sum(1 for ln in tile_asm if ln.strip().startswith("v_mfma_f32_16x16x32_f16")) > 0


def test_no_inline_vector_instructions_in_tile_kernel(tile_asm):
    inside, found = False, []
    for ln in tile_asm:
        s = ln.strip()
        if s.startswith(";;#ASMSTART"):
            inside = True
        elif s.startswith(";;#ASMEND"):
            inside = False
        elif inside and re.match(r"^v_", s):
            found.append(s)
    assert not found, found[:5]
