"""Writes tests/golden/ref_vectors.npz: outputs of the reference's own CPU
solvers (oracle/_ref, compiled from the reference sources by
oracle/Makefile.ref) on small seeded inputs.

    make -C oracle -f Makefile.ref && python tests/golden/make_golden.py

Inputs are rebuilt from filtered_tdma_amd/cases.py (seeded, counter-based);
the file stores each case's parameters, a SHA-256 of the inputs (so a change
of the generator is caught) and the reference's solved D.

Two families of cases:
  backend : TdmaBackend (apps/heat_cpu/tdma_backend.cpp) under mpiexec -n P,
            CPU semantics, set_rho from column 0, eps_constant;
  steps   : the Filtered step functions with a forced J per partition
            (filtered_tdma_kernels.hpp), whose J-dependent sweeps equal the
            GPU kernels' for J in [1, m-2].
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(HERE, "ref_vectors.npz")


def inputs(case):
    from filtered_tdma_amd import cases

    if case["gen"] == "general":
        A, B, C, D = cases.general_batch(case["n_sys"], case["n_row"], seed=case["seed"])
    else:
        A, B, C, D, _ = cases.toeplitz_batch(case["n_sys"], case["n_row"], case["r_lo"],
                                             case["r_hi"], seed=case["seed"])
    return A, B, C, D


def digest(arrs):
    h = hashlib.sha256()
    for a in arrs:
        h.update(np.ascontiguousarray(a, dtype="<f8").tobytes())
    return h.hexdigest()


def case_list():
    out = []
    for P in (1, 2, 3, 4):
        for cyc in (0, 1):
            out.append(dict(mode="backend", kind="pascal", gen="general", n_sys=16, n_row=61,
                            seed=31, P=P, cyclic=cyc, do_rho=0, eps_constant=0.0))
    for kind in ("filtered", "filtered_v2"):
        for P in (1, 2, 4):
            for cyc in (0, 1):
                for (rlo, rhi, epsc) in ((0.05, 0.3, 1e-6), (0.5, 20.0, 1e-3)):
                    out.append(dict(mode="backend", kind=kind, gen="toeplitz", n_sys=12,
                                    n_row=96, seed=5, r_lo=rlo, r_hi=rhi, P=P, cyclic=cyc,
                                    do_rho=1, eps_constant=epsc))
    for kind in ("filtered", "filtered_v2"):
        for P, J in ((2, 1), (2, 7), (3, 2), (3, 29), (4, 22)):
            for cyc in (0, 1):
                out.append(dict(mode="steps", kind=kind, gen="toeplitz", n_sys=8, n_row=96,
                                seed=8, r_lo=0.1, r_hi=0.45, P=P, cyclic=cyc, J=J))
    return out


def main():
    import oracle as O

    if not O.ref_available():
        sys.exit("reference CPU build missing: make -C oracle -f Makefile.ref")
    meta, arrays = [], {}
    for k, case in enumerate(case_list()):
        A, B, C, D = inputs(case)
        case = dict(case, sha256=digest((A, B, C, D)))
        x = O.ref_run(case["mode"], case["kind"], A, B, C, D, case["P"],
                      cyclic=bool(case["cyclic"]), do_rho=bool(case.get("do_rho", 0)),
                      eps_constant=case.get("eps_constant", 0.0), J=case.get("J"))
        arrays[f"x{k}"] = x
        meta.append(case)
        print(k, case["mode"], case["kind"], "P", case["P"], "cyclic", case["cyclic"], flush=True)
    np.savez_compressed(OUT, meta=np.array(json.dumps(meta)), **arrays)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
