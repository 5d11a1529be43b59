"""Build ``libhipnufft.so`` for gfx950 with hipcc (in-tree, no JIT cache).

Usage: ``python -m hipnufft._build`` (from ``mri-nufft_amd/``) or
``hipnufft._build.build()``.  Objects go to ``mri-nufft_amd/_build/`` and the
library to ``hipnufft/_lib/libhipnufft.so``.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent  # mri-nufft_amd/
REPO = ROOT.parent
CSRC = ROOT / "csrc"
INCLUDE = REPO / "include"
OBJ = ROOT / "_build"
LIB = PKG / "_lib" / "libhipnufft.so"
SOURCES = ["sort.hip", "spread_interp.hip", "fft.hip", "maps.hip", "orc.hip", "capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HIPNUFFT_ARCH", "gfx950")

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return res


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every HIP source and link the shared library; returns its path."""
    OBJ.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    objs = [OBJ / (Path(s).stem + ".o") for s in SOURCES]

    def compile_one(i):
        cmd = [HIPCC, *CFLAGS, "-c", str(CSRC / SOURCES[i]), "-o", str(objs[i])]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)

    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(compile_one, range(len(SOURCES))))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
