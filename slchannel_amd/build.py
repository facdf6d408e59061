"""Build libslc_hip.so (gfx950) in-tree with hipcc.

    python -m slchannel_amd.build        # or __graft_entry__.build()

No CUDA, no hipify, no JIT cache: the shared library lands next to this file
(slchannel_amd/libslc_hip.so) so it ships with the source tree.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libslc_hip.so")
SOURCES = ["slc_api.hip", "slc_sl.hip", "slc_ops.hip", "slc_tiled.hip", "slc_step.hip"]
ARCH = os.environ.get("SLC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
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


def build(verbose=False):
    """Compile every source into slchannel_amd/libslc_hip.so (always rebuilds)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *srcs, "-o", tmp, "-lrocfft"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) building {OUT}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
