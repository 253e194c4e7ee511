"""Builds libadeco_amd.so (hipcc, gfx950) in-tree.

Every HIP source under adeco_amd/csrc is compiled for gfx950 with packed-FP32 instructions
disabled (-packed-fp32-ops) and hardware FP64 atomics enabled; the library is written next to
this file so it travels with the source tree, and the adeco_resample program is linked against it.
"""

from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libadeco_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    "-ffp-contract=fast",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]
LIBS = ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx"]  # roctx ranges for rocprofv3 --marker-trace


def sources():
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def build_cli(verbose: bool = False) -> str:
    """Links the adeco_resample program (adeco_amd/cpp/adeco_resample.cpp) against the library."""
    exe = os.path.join(HERE, "adeco_resample")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(HERE, "cpp", "adeco_resample.cpp"), f"-L{HERE}", "-ladeco_amd", "-Wl,-rpath,$ORIGIN",
           "-o", exe + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("adeco_resample link failed:\n" + res.stdout + res.stderr)
    os.replace(exe + ".tmp", exe)
    return exe


def build(verbose: bool = False) -> str:
    srcs = sources()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *srcs, *LIBS, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = (res.stdout or "") + (res.stderr or "")
    # hipcc's host pass does not know the packed-fp32 feature and says so; that line is expected
    noise = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(noise))
    if verbose and noise:
        print("\n".join(noise))
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
