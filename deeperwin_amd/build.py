"""Build the gfx950 shared library ``deeperwin_amd/libdwmi.so`` from ``deeperwin_amd/csrc`` with hipcc.

Every HIP translation unit is compiled with ``--offload-arch=gfx950`` and with the packed-FP32 target
feature switched off; the objects are linked into one C-ABI library that the Python host loads through
ctypes.  ``python -m deeperwin_amd.build`` rebuilds it in place.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libdwmi.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["gemm.hip", "wf_kernels.hip", "api.cpp", "observables.hip", "grad.hip", "threefry.hip", "proposals.hip",
           "dpe4.hip", "dpe4_grad.hip", "noalloc.hip"]
COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def _compile(src: str, objdir: Path) -> Path:
    obj = objdir / (src + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(CSRC / src), "-o", str(obj)]
    if src.endswith(".hip"):
        cmd.insert(1, "-xhip")
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, objdir), SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
