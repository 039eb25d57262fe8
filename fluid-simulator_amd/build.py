"""Builds libpbf_amd.so for gfx950 with hipcc, in-tree.

HIP sources (*.hip) are compiled for the device and the host; host-only
sources (host/*.cpp) are compiled as C++ against the HIP runtime headers.
The shared library lands in fluid-simulator_amd/pbf_amd/libpbf_amd.so.

Usage: python fluid-simulator_amd/build.py [--jobs N] [--verbose]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "pbf_amd" / "libpbf_amd.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          f"-I{REPO / 'include'}", f"-I{CSRC}", f"-I{CSRC / 'host'}"]
# Packed FP32 instructions stay disabled on the device (they are emitted by
# the compiler on its own otherwise).
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fno-gpu-rdc"]


# Host-only sources: C++ against the HIP runtime headers, no device pass.
HOST = ["-x", "c++", "-D__HIP_PLATFORM_AMD__=1", "-I/opt/rocm/include"]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip")) + sorted((CSRC / "host").glob("*.cpp"))


def compile_one(src: Path, verbose: bool) -> Path:
    rel = src.relative_to(CSRC)
    obj = BUILD / (str(rel).replace(os.sep, "_") + ".o")
    cmd = [HIPCC, *COMMON]
    if src.suffix == ".hip":
        cmd += DEVICE
    else:
        cmd += HOST
    cmd += ["-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{res.stdout}\n{res.stderr}")
    err = "\n".join(l for l in res.stderr.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if err and verbose:
        print(err, file=sys.stderr)
    return obj


def build(jobs: int = 8, verbose: bool = False) -> Path:
    if not Path(HIPCC).exists() and shutil.which(HIPCC) is None:
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    lib = build(a.jobs, a.verbose)
    print(f"built {lib}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
