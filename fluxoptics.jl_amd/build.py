"""Build libfluxhip.so (gfx950) in-tree with hipcc.

Usage: python fluxoptics.jl_amd/build.py [--verbose]
The library lands next to the Python host layer (fluxoptics/libfluxhip.so), where the
loader looks for it.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "fluxoptics", "libfluxhip.so")
SOURCES = ["fluxhip.hip"]
HEADERS = ["fh_dft.hpp", "fh_lines.hpp", "fh_ops.hpp", "fh_bpm.hpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off: plain float code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libfluxhip")


def build(verbose: bool = False) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, *[os.path.join(CSRC, s) for s in SOURCES], "-o", OUT + ".tmp"]
    t0 = time.time()
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}")
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        msgs = [l for l in (res.stdout + res.stderr).splitlines()
                if "packed-fp32-ops" not in l and l.strip()]
        if msgs:
            print("\n".join(msgs))
        print(f"built {OUT} in {time.time() - t0:.1f}s")
    return OUT


if __name__ == "__main__":
    build(verbose="--verbose" in sys.argv or "-v" in sys.argv)
