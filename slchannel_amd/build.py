"""Build libslc_hip.so (gfx950) in-tree with hipcc.

    python -m slchannel_amd.build        # or __graft_entry__.build()

No CUDA, no hipify, no JIT cache: the shared library lands next to this file
(slchannel_amd/libslc_hip.so) so it ships with the source tree.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libslc_hip.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["slc_api.hip", "slc_sl.hip", "slc_ops.hip", "slc_tiled.hip", "slc_step.hip",
           "slc_ens.hip", "slc_ens_kernels.hip", "slc_ens_tile_d4.hip", "slc_ens_tile_d6.hip",
           "slc_ens_tile_f4.hip", "slc_ens_tile_f6.hip", "slc_stats.hip",
           "slc_diag.hip", "slc_imex.hip", "slc_init.hip"]
ARCH = os.environ.get("SLC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # feet and stencil indices must be the IEEE results of the written
    # operation order; fused multiply-adds only where written as fma()
    "-ffp-contract=off",
    # no packed-fp32 instructions anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd[-3:])}")


def build(verbose=False, jobs=None):
    """Compile every source (in parallel) and link slchannel_amd/libslc_hip.so.
    Always rebuilds."""
    os.makedirs(BUILD, exist_ok=True)
    objs = [os.path.join(BUILD, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, s), "-o", o]
            for s, o in zip(SOURCES, objs)]
    jobs = jobs or min(len(cmds), max(1, min(os.cpu_count() or 1, 8)))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = OUT + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lrocfft"],
         verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
