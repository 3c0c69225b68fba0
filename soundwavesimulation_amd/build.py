"""Compile ``csrc/swfd_kernels.hip`` into ``libswfd.so`` for gfx950.

``python -m soundwavesimulation_amd.build`` (or ``__graft_entry__.build()``)
runs one ``hipcc`` command; the library lands next to this file so it ships
with the source tree. Packed-FP32 instruction selection is disabled and FMA
contraction is off (the kernels place their fused multiply-adds explicitly).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SOURCE = os.path.join(HERE, "csrc", "swfd_kernels.hip")
LIBRARY = os.path.join(HERE, "libswfd.so")
ARCH = os.environ.get("SWFD_OFFLOAD_ARCH", "gfx950")


def hipcc_command(output=LIBRARY, arch=ARCH):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return [
        hipcc, f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(REPO, "include"),
        "-o", output, SOURCE,
    ]


def build(verbose=True):
    """Compile the library (always recompiles) and return its path."""
    tmp = LIBRARY + ".tmp"
    cmd = hipcc_command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    build()
    sys.exit(0)
