"""Build recipe of libftdma.so (hipcc, gfx950), written in-tree next to this file.

    python -m filtered_tdma_amd.build            # build
    python -m filtered_tdma_amd.build --resource # also print per-kernel VGPR/LDS use
    python -m filtered_tdma_amd.build --incremental  # only sources changed since the last build

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


def _filtered(out: str) -> str:
    # hipcc's host pass warns that it does not know the packed-fp32 feature
    return "\n".join(ln for ln in out.splitlines() if "packed-fp32-ops" not in ln)


def _stale(src: str, obj: str) -> bool:
    if not os.path.exists(obj):
        return True
    deps = [src] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")] + \
        [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    return max(os.path.getmtime(d) for d in deps) > os.path.getmtime(obj)


def build(resource_usage: bool = False, verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every source to an object in parallel (one hipcc each), then
    links libftdma.so. incremental=True recompiles only the sources (or
    headers) changed since their object was built."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(os.path.basename(src))[0] + ".o")
        if incremental and not resource_usage and not _stale(src, obj):
            procs.append((obj, None))
            continue
        cmd = [HIPCC, *compile_flags, "-c", "-x", "hip", src, "-o", obj]
        if resource_usage:
            cmd.insert(1, "-Rpass-analysis=kernel-resource-usage")
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    failed, objs = [], []
    for obj, pr in procs:
        if pr is None:
            objs.append(obj)
            continue
        out, _ = pr.communicate()
        out = _filtered(out)
        if pr.returncode != 0:
            failed.append(out)
        elif resource_usage or verbose:
            print(out)
        objs.append(obj)
    if failed:
        sys.stderr.write("\n".join(failed) + "\n")
        raise RuntimeError("hipcc failed")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-lhipfft",
           "-o", LIB + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(_filtered(res.stdout + res.stderr) + "\n")
        raise RuntimeError(f"link failed with exit code {res.returncode}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(resource_usage="--resource" in sys.argv, verbose="-v" in sys.argv,
          incremental="--incremental" in sys.argv)
    print(LIB)
