"""Build the native parts in-tree.

* ``merlict_development_kit_amd/libmlgpu.so``: host scene code (C++) and the
  HIP kernels for gfx950, one shared library exporting the C ABI of
  ``include/mlgpu.h``.
* ``oracle/liboracle.so`` (and, when the reference sources are present,
  ``oracle/_ref``): built by ``oracle/Makefile``; test infrastructure only.

Run ``python -m merlict_development_kit_amd.build`` or call :func:`build`.
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
LIB = PKG / "libmlgpu.so"
ARCH = os.environ.get("MLGPU_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    # FP64 leaf solves follow the reference's operation order; no contraction
    # (the FP32 traversal writes its FMAs explicitly)
    "-ffp-contract=off",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]

SOURCES = ["scene_desc.cpp", "builders.cpp", "scene_build.cpp", "api.cpp", "propagate.hip", "sources.hip", "arrivals.hip", "corsika.hip", "surrogate.hip",
           "surrogate_api.cpp", "signal.hip", "signal_api.cpp", "calibration.hip", "render.hip", "pool.cpp",
           "eventio.cpp", "event_tape.cpp"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed: {' '.join(map(str, cmd))}")
    return proc.stdout


def build_library(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    objdir = ROOT / "_build" / "mlgpu"
    objdir.mkdir(parents=True, exist_ok=True)
    def compile_one(src: str) -> str:
        obj = objdir / (src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        out = _run([hipcc, *HIP_FLAGS, *lang, "-c", str(CSRC / src), "-o", str(obj)])
        if verbose and out.strip():
            print(out)
        return str(obj)

    # the translation units are independent: compile them side by side
    workers = max(1, min(8, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        objs = list(pool.map(compile_one, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> None:
    """Build oracle/liboracle.so and, when /root/reference exists, oracle/_ref:
    the reference build, and the reference-side GPU binding (facade/) with its
    check against the reference (needs libmlgpu.so, so after build_library)."""
    _run(["make", "-C", str(ROOT / "oracle"), "-j8", "oracle"])
    if Path("/root/reference").is_dir():
        _run(["make", "-C", str(ROOT / "oracle"), "-j8", "ref"])
        _run(["make", "-C", str(ROOT / "oracle"), "facade"])


def build(verbose: bool = False) -> None:
    build_library(verbose=verbose)
    build_oracle()


if __name__ == "__main__":
    build(verbose=True)
    print(f"built {LIB}")
