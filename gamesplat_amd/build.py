"""Build libgamesplat_hip.so (gfx950) in-tree with hipcc.

Usage: python -m gamesplat_amd.build [--jobs N]

Every HIP source is compiled for gfx950 with
  -ffp-contract=off                     every float operation stays a separate
                                        IEEE operation in source order
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                        no packed FP32 instructions
and linked with the host sources into gamesplat_amd/libgamesplat_hip.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libgamesplat_hip.so"
HEADLESS = PKG / "gs_headless"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GS_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _sources():
    hip = sorted(CSRC.glob("*.hip"))
    cpp = sorted(CSRC.glob("*.cpp"))
    return hip, cpp


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-I", str(CSRC), *DEVICE]
    if src.suffix != ".hip":
        cmd += ["-x", "hip"]
    cmd += ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    hip, cpp = _sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, hip + cpp))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *DEVICE, "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # headless renderer driver (C++ host code over the C ABI)
    rsrc = sorted((PKG / "renderer").glob("*.cpp"))
    cmd = [HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-o", str(HEADLESS), *map(str, rsrc), "-L", str(PKG),
           "-lgamesplat_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"headless driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB} and {HEADLESS}")
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args(argv)
    build(args.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
