"""Build the native library `libnanodrr_amd.so` in-tree with hipcc for gfx950.

Usage: ``python -m nanodrr_amd.build`` (or ``__graft_entry__.build()``).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SOURCES = [PKG / "csrc" / name for name in ("raymarch.hip", "camera.hip", "register.hip", "conv.hip", "head.hip")]
OUTPUT = PKG / "libnanodrr_amd.so"

# -packed-fp32-ops: keep the backend from forming packed fp32 instructions.
# -amdgpu-mfma-vgpr-form: matrix-core accumulators in ordinary VGPRs; in the accumulator
# registers the allocator rotated each K loop's sums through copies (up to 76 moves per 32 MFMAs).
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; the HIP toolchain (ROCm) is required to build nanodrr_amd")
    return path


def build(verbose: bool = False, extra_flags: list[str] | None = None) -> Path:
    """Compile every HIP source into OUTPUT (always recompiles)."""
    tmp = OUTPUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, *(extra_flags or []), f"-I{ROOT / 'include'}",
           *map(str, SOURCES), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    out = build(verbose="-v" in sys.argv)
    print(out)
