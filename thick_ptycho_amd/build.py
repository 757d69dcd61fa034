"""Build the in-tree HIP library ``thick_ptycho_amd/libtp_hip.so`` for gfx950.

Usage: ``python -m thick_ptycho_amd.build`` (or ``__graft_entry__.build()``).
Each source is compiled to an object with hipcc, then linked into one shared
library that exports the C ABI declared in ``include/thick_ptycho_amd.h``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtp_hip.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["tp_status.cpp", "tp_tri.hip", "tp_tri_sweep_c16.hip", "tp_tri_sweep_c32.hip", "tp_tri_sweep_c8.hip",
           "tp_tri_sweep_c4.hip", "tp_tri_sweep_c2.hip", "tp_tri_sweep_c1.hip", "tp_triw.hip", "tp_ms.hip", "tp_fftany.hip", "tp_band.hip", "tp_learn.hip", "tp_krylov.hip", "tp_ms2d.hip",
           "tp_pint.hip", "tp_pie.hip", "tp_window.hip", "tp_filter.hip", "tp_refine.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-I" + os.path.join(ROOT, "include"),
    # packed FP32 (v_pk_*_f32) stays off for every kernel in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") or src.endswith(".cpp") else []
    cmd = [HIPCC, *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
