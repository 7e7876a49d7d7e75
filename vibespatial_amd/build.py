"""Build libvsjoin.so from the HIP sources in csrc/ for gfx950.

    python -m vibespatial_amd.build

hipcc cross-compiles without a GPU.  Each source is compiled to an object in
_build/ (in parallel), then the objects are linked into the shared library.
Flags:
  -ffp-contract=off   the exact-arithmetic transforms and the FP32 error bounds
                      assume separately rounded operations (no FMA contraction)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions in generated code
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvsjoin.so")
OBJ_DIR = os.path.join(os.path.dirname(HERE), "_build")
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


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}: {' '.join(cmd[-3:])}")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source and link libvsjoin.so (always recompiles)."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[hipcc(), *FLAGS, "-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
    jobs = jobs or min(len(cmds), max(1, min(8, len(os.sched_getaffinity(0)))))
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
