"""Build the gfx950 HIP library ``gbgpu_amd/libgbgpu_hip.so`` in-tree.

Usage: ``python -m gbgpu_amd.build`` (or ``__graft_entry__.build()``).
Each ``csrc/*.hip`` file is compiled to an object in parallel, then linked
into one shared library exposing the C ABI declared in ``include/gbgpu_hip.h``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libgbgpu_hip.so"
ARCH = os.environ.get("GBGPU_HIP_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 VALU forms are kept out of the device code
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    # matrix-core results in architectural VGPRs rather than accumulation
    # registers: the transform's accumulators then share the kernel's VGPR
    # budget instead of raising its register total (occupancy 7, not 5)
    "-mllvm",
    "-amdgpu-mfma-vgpr-form=1",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _compile(src: Path, obj: Path) -> None:
    cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr}")


def build(verbose: bool = True) -> Path:
    srcs = sorted(CSRC.glob("*.hip"))
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = [objdir / (s.stem + ".o") for s in srcs]
    # every file at once: the three largest take most of the time
    with ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        list(ex.map(lambda so: _compile(*so), zip(srcs, objs)))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
