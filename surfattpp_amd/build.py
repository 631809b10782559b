"""Builds libsurfatt_eik.so (gfx950) in-tree with hipcc.

Every translation unit under csrc/ is compiled in its own hipcc process (in
parallel), then linked into one shared library.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsurfatt_eik.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["sat_eik.hip", "k_pair64.hip", "k_pair32.hip", "k_phase64.hip", "k_phase32.hip", "k_mw64.hip",
           "k_mw32.hip", "k_pair2.hip", "k_wide64.hip", "k_wide32.hip", "k_widefused64.hip",
           "k_widefused32.hip", "sat_combine.hip", "sat_topo.hip", "sat_surfdisp.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    # the FP64 path reproduces the reference's operation order: no contraction
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(src: str, extra: list[str], verbose: bool) -> tuple[str, subprocess.CompletedProcess]:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *extra, "-c", "-o", obj, os.path.join(CSRC, src)]
    if verbose:
        print(" ".join(cmd), flush=True)
    return obj, subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    """Compile and link; `extra` flags go to every hipcc compile."""
    os.makedirs(OBJ, exist_ok=True)
    extra = list(extra or [])
    jobs = min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, extra, verbose), SOURCES))
    failed = False
    for obj, r in results:
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            failed = True
        elif verbose and (r.stdout or r.stderr):
            sys.stderr.write(r.stdout + r.stderr)
    if failed:
        raise RuntimeError("hipcc failed building libsurfatt_eik.so")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[o for o, _ in results]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libsurfatt_eik.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
