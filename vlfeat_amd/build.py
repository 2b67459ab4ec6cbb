"""Build recipe of the MI355X VLFeat-compatible library (``libvl.so``).

Host code is C99 compiled with gcc; device code is HIP C++ compiled with
hipcc for gfx950 only.  Everything links into one shared object that exports
the VLFeat C API (``include/vl/*.h``) and the device-level ABI
(``include/vlhip.h``).  The library is written in-tree to
``vlfeat_amd/lib/libvl.so`` so it travels with the repository snapshot.

Usage: ``python -m vlfeat_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
HOST_SRC = PKG / "csrc" / "host"
HIP_SRC = PKG / "csrc" / "hip"
BUILD = PKG / "_build"
LIB = PKG / "lib" / "libvl.so"
SYNTH_SRC = PKG / "csrc" / "synth" / "vlsynth.c"
SYNTH_LIB = PKG / "lib" / "libvlsynth.so"

OFFLOAD_ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={OFFLOAD_ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CC_FLAGS = [
    "-std=gnu99", "-O2", "-fPIC", "-fvisibility=hidden", "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-pthread",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the library's kernels must be compiled with ROCm's hipcc")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        msg = "\n".join(
            line for line in (res.stdout + res.stderr).splitlines()
            if "packed-fp32-ops" not in line
        )
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{msg}")


def build(verbose: bool = False) -> Path:
    """Compile every source and link ``libvl.so``; returns its path."""
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    for src in sorted(HIP_SRC.glob("*.hip")):
        obj = BUILD / (src.stem + ".hip.o")
        jobs.append([hipcc, *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{HIP_SRC}", "-c", str(src), "-o", str(obj)])
    for src in sorted(HOST_SRC.glob("*.c")):
        obj = BUILD / (src.stem + ".c.o")
        jobs.append(["gcc", *CC_FLAGS, "-fopenmp", f"-I{INCLUDE}", f"-I{HOST_SRC}", "-c", str(src), "-o", str(obj)])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for cmd, fut in [(c, ex.submit(_run, c)) for c in jobs]:
            fut.result()
            if verbose:
                print(" ".join(cmd))
    objs = [j[-1] for j in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs,
          "-Wl,-Bsymbolic", "-lgomp", "-lpthread", "-lm"])
    os.replace(tmp, LIB)
    build_synth()
    return LIB


def build_synth() -> Path:
    """The seeded data generator of the benchmark workloads (host C, OpenMP):
    ``libvlsynth.so`` beside ``libvl.so``."""
    LIB.parent.mkdir(exist_ok=True)
    tmp = SYNTH_LIB.with_suffix(".so.tmp")
    _run(["gcc", *CC_FLAGS, "-fopenmp", f"-I{INCLUDE}", "-shared", str(SYNTH_SRC), "-o", str(tmp), "-lm"])
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
