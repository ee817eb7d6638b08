"""Builds the native parts in-tree.

* ``hdcodex_amd/libhdcx.so`` — the product: C++ host code and HIP kernels for
  gfx950 behind the C ABI declared in ``include/hdcx.h``.
* ``oracle/liboracle.so`` — the CPU oracle (test infrastructure only), built
  with ``-ffp-contract=off`` so that its FP32 arithmetic is the IEEE sequence
  the restatement spells out.
* ``oracle/liboracle_x86-64-v3.so`` and ``oracle/liboracle_x86-64-v4.so`` —
  the same source at ``-O3`` for the AVX2 and AVX-512 levels of x86-64: the
  CPU baseline loads the highest level the host supports (the library is
  built here, before the run, so ``-march=native`` would describe this
  machine rather than the host that runs the baseline). Contraction stays
  off, so results equal the checker's.

Usage: ``python -m hdcodex_amd.build [--lib] [--oracle]`` (default: both).
"""

from __future__ import annotations

import argparse
import concurrent.futures as futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "hdcodex_amd"
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libhdcx.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

ARCH = os.environ.get("HDCX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are kept out of the generated code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"),
]

HOST_SOURCES = ["packer.cpp", "procedural.cpp", "procedural_showcase.cpp", "procedural_mixed.cpp", "bvh_host.cpp",
                "lbvh_host.cpp"]
HIP_SOURCES = ["kernels.hip", "shade.hip", "rcache.hip", "lbvh.hip", "tables.hip", "scene_store.cpp", "tracer.cpp",
               "batch.cpp"]
# The shading stages divide and take square roots at the hardware rate
# (within 2.5 ulp); the other files keep IEEE division (bit-exact hits).
EXTRA_FLAGS = {"shade.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt"]}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")


def build_library() -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    jobs = []
    objects = []
    for src in HOST_SOURCES:
        obj = BUILD / (src + ".o")
        objects.append(obj)
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", str(ROOT / "include"), "-c",
                     str(CSRC / src), "-o", str(obj)])
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        objects.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, *EXTRA_FLAGS.get(src, []), "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)])
    with futures.ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 4)) as pool:
        list(pool.map(_run, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects)])
    os.replace(tmp, LIB)
    return LIB


ORACLE_LEVELS = ("x86-64-v3", "x86-64-v4")


def oracle_variant(level: str) -> Path:
    return ORACLE_DIR / f"liboracle_{level}.so"


def build_oracle() -> Path:
    jobs = [(ORACLE_LIB, ["-O2"])] + [(oracle_variant(level), ["-O3", f"-march={level}", "-mtune=generic"])
                                        for level in ORACLE_LEVELS]

    def one(job):
        out, flags = job
        tmp = out.with_suffix(".so.tmp")
        _run(["g++", *flags, "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-pthread",
              "-I", str(ROOT / "include"), str(ORACLE_DIR / "oracle.cpp"), "-o", str(tmp)])
        os.replace(tmp, out)

    with futures.ThreadPoolExecutor(max_workers=len(jobs)) as pool:
        list(pool.map(one, jobs))
    return ORACLE_LIB


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    args = ap.parse_args(argv)
    both = not (args.lib or args.oracle)
    if both or args.oracle:
        print("built", build_oracle())
    if both or args.lib:
        print("built", build_library())
    return 0


if __name__ == "__main__":
    sys.exit(main())
