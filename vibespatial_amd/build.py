"""Build libvsjoin.so from the HIP sources in csrc/ for gfx950.

    python -m vibespatial_amd.build

hipcc cross-compiles without a GPU.  Flags:
  -ffp-contract=off   the exact-arithmetic transforms and the FP32 error bounds
                      assume separately rounded operations (no FMA contraction)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions in generated code
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvsjoin.so")
ARCH = os.environ.get("VSJOIN_ARCH", "gfx950")

SOURCES = [
    "vs_api.hip", "vs_sort.hip", "vs_index.hip", "vs_join.hip", "vs_finish.hip", "vs_refine.hip", "vs_learned.hip",
    "vs_batch.hip", "vs_wkb.hip", "vs_nearest.hip",
]

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    """Compile every source into libvsjoin.so (always recompiles)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *FLAGS, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
