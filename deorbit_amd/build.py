"""Build the gfx950 shared library ``libdsph.so`` in-tree with hipcc.

The library is one translation unit (csrc/dsph.hip including the kernel
headers).  Packed-FP32 instruction selection is disabled for all device code
(``-target-feature -packed-fp32-ops``); everything else is plain -O3.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libdsph.so")
ARCH = os.environ.get("DSPH_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def compile_command(out: str = LIB, extra: list[str] | None = None) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            f"-I{INCLUDE}", f"-I{CSRC}", os.path.join(CSRC, "dsph.hip"), "-o", out] + (extra or [])


def build_library(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = compile_command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libdsph.so")
    if verbose and r.stderr.strip():
        # hipcc's host pass reports the device-only target feature; keep other warnings visible
        lines = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if any(ln.strip() for ln in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    os.replace(tmp, LIB)
    return LIB


CLI_SRC = os.path.join(HERE, "cli", "dsph_gpu_ic.cpp")
CLI = os.path.join(HERE, "bin", "dsph_gpu_ic")


def build_cli(verbose: bool = False) -> str:
    """The gpu_ic-compatible driver: host C++ over the C ABI, linked to libdsph.so."""
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", CLI_SRC, "-o", CLI, f"-L{HERE}", "-ldsph",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building dsph_gpu_ic")
    return CLI


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_cli(verbose=True))
