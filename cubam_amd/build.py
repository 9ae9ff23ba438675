"""Build libcubam_amd.so (hipcc, gfx950) in the package directory."""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libcubam_amd.so"
ARCH = os.environ.get("CUBAM_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-shared",
    "-fPIC",
    # no packed FP32 instructions anywhere in this library
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.h")) + [HERE.parent / "include" / "cubam_amd.h"]


def build(force: bool = True, verbose: bool = False) -> Path:
    """Compile the HIP library (always when ``force``, else only when a source is newer)."""
    if not force and LIB.exists():
        newest = max(p.stat().st_mtime for p in sources())
        if LIB.stat().st_mtime >= newest:
            return LIB
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB.with_name(LIB.name + f".tmp{os.getpid()}")
    cmd = [hipcc, *HIPCC_FLAGS, "-o", str(tmp), str(CSRC / "cubam_amd.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
