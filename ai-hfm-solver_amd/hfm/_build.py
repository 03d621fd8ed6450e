"""Build libhfm.so for gfx950 with hipcc (in-tree, next to this file)."""
from __future__ import annotations

import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(PKG_DIR), "csrc")
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(PKG_DIR)), "include")
LIB_PATH = os.path.join(PKG_DIR, "libhfm.so")
SOURCES = ("hfm_ref.hip", "hfm_mg.hip", "hfm_flow2d.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 instructions are disabled for every device pass.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    """Compile every HIP source into ``libhfm.so``; returns its path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) building libhfm.so")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_asan_host(out: str, verbose: bool = False) -> str:
    """The same library with its host code built under AddressSanitizer (the
    device code as usual, at -O1), for the C-ABI argument-check tests: load it
    into a process that preloads the compiler's ASan runtime."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    flags = [f for f in HIP_FLAGS if f != "-O3"] + [
        "-O1", "-g", "-Xarch_host", "-fsanitize=address",
        "-Xarch_host", "-fno-omit-frame-pointer"]
    cmd = [HIPCC, *flags, f"-I{INCLUDE}", f"-I{CSRC}", "-o", out, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) building {out}")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
