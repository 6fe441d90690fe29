"""Build recipe of the HIP library (hipcc, gfx950), in-tree.

    python -m landschaft_amd.build

compiles landschaft_amd/csrc into landschaft_amd/_lib/liblandschaft_circuit.so.
Every HIP translation unit is built with packed FP32 instructions disabled.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "_lib"
LIB = OUT / "liblandschaft_circuit.so"
ROOT = PKG.parent
ARCH = os.environ.get("LSC_OFFLOAD_ARCH", "gfx950")

SOURCES = ["capi.hip", "omniscape.hip", "system.hip", "circuitscape.hip", "plan.cpp", "grid_text.cpp"]
# MFMA accumulators in ordinary vector registers (gfx950 allows either file):
# no copies through the accumulation registers around every tile's MFMAs
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form=true", "-Wall", "-Wno-unused-result", "-I",
         str(ROOT / "include")]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build landschaft_amd)")


def _compile(src: Path, obj: Path) -> None:
    cmd = [hipcc(), *FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stdout}\n{r.stderr}")


def build(verbose: bool = False) -> Path:
    """Compile every source (always) and link the shared library."""
    obj_dir = OUT / "_build"
    obj_dir.mkdir(parents=True, exist_ok=True)
    sources = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    objs = [obj_dir / (s.stem + ("_host" if s.suffix == ".cpp" else "") + ".o") for s in sources]
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        list(ex.map(lambda so: _compile(*so), zip(sources, objs)))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
