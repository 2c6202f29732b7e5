"""Build recipe of libftdma.so (hipcc, gfx950), written in-tree next to this file.

    python -m filtered_tdma_amd.build            # build
    python -m filtered_tdma_amd.build --resource # also print per-kernel VGPR/LDS use

Flags:
  -ffp-contract=off      no product is fused into a sum: results equal the
                         CPU restatement operation for operation;
  -packed-fp32-ops off   no packed FP32 instructions are generated.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libftdma.so")
SOURCES = [os.path.join(CSRC, "ftdma_kernels.hip"), os.path.join(CSRC, "ftdma_api.cpp"),
           os.path.join(CSRC, "ftdma_heat.hip"), os.path.join(CSRC, "ftdma_batched.hip"),
           os.path.join(CSRC, "ftdma_poisson.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-I", os.path.join(ROOT, "include"),
]


def build(resource_usage: bool = False, verbose: bool = False) -> str:
    srcs = []
    for src in SOURCES:
        srcs += ["-x", "hip", src]
    cmd = [HIPCC, *FLAGS, *srcs, "-lhipfft", "-o", LIB + ".tmp"]
    if resource_usage:
        cmd.insert(1, "-Rpass-analysis=kernel-resource-usage")
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = res.stdout + res.stderr
    # hipcc's host pass warns that it does not know the packed-fp32 feature
    lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
    if res.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    if resource_usage or verbose:
        print("\n".join(lines))
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(resource_usage="--resource" in sys.argv, verbose="-v" in sys.argv)
    print(LIB)
