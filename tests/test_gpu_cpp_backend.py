"""The C++ boundary as a reference caller uses it: a program built against
include/tdma_backend_gpu.hpp and libftdma.so (tests/cpp/backend_check.cpp)
runs the reference's per-step protocol (set_eps_constant, set_rho_device,
solve / solve_cyclic, last_J, last_gpu_ms, last_comm_ms;
apps/heat_gpu/solve_theta.cu:522-552) for every kind, with the class's
extensions (set_guard, set_j_mode, last_J_systems) on some cases, and
captures the first solve of a fresh handle into a hipGraph.

Checked here: x bit-identical to the oracle with the engine's partitioning,
J equal to the oracle's, last_gpu_ms > 0 by default (the reference's solve
always profiles, tdma_backend_gpu.cu:35-46) and 0 after
set_profiling(false), last_comm_ms 0, the parse table, and three graph
replays bit-identical to the eager solve (the capture runs in global mode,
where an allocation inside the solve would fail it).
"""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from filtered_tdma_amd import Kind, cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "cpp", "backend_check")

GUARD, LEARNED = 1, 2

# (kind, n_sys, n_row, partitions, cyclic, eps_constant, options)
CASES = [
    (Kind.PASCAL, 256, 128, 4, 0, 0.0, 0),
    (Kind.PASCAL, 100, 60, 1, 1, 0.0, 0),
    (Kind.FILTERED_V2, 256, 128, 4, 0, 1e-3, 0),
    (Kind.FILTERED, 256, 128, 4, 0, 1e-3, 0),
    (Kind.FILTERED, 96, 60, 3, 1, 1e-2, 0),
    (Kind.FILTERED_V2, 130, 200, 2, 0, 1e-4, 0),
    (Kind.FILTERED_V2, 256, 128, 4, 0, 1e-3, GUARD),
    (Kind.FILTERED, 96, 60, 3, 1, 1e-2, GUARD),            # periodic, guarded
    (Kind.FILTERED, 256, 128, 4, 0, 1e-6, LEARNED),
    (Kind.FILTERED, 130, 300, 2, 0, 1e-6, LEARNED | GUARD),  # general kernels
]


def oracle_solution(kind, A, B, C, D, P, cyclic, eps_constant, options):
    """x, the per-partition J (None where it is not checked) and the records
    (None without options)."""
    if kind == Kind.PASCAL:
        if P == 1:
            return O.thomas(A, B, C, D, sem=O.SEM_GPU, cyclic=bool(cyclic)), None, None
        return O.pascal(A, B, C, D, P, sem=O.SEM_GPU, cyclic=bool(cyclic)), None, None
    version = 2 if kind == Kind.FILTERED_V2 else 1
    n_sys = D.shape[1]
    fb = np.zeros((P, n_sys), dtype=np.int32)
    if options & LEARNED:
        x, Jr = O.filtered_learned(A, B, C, D, P, eps_constant, cyclic=bool(cyclic))
        Jsys, fb, Jp = Jr & 0xFFFF, (Jr >> 16) & 1, None
        if not options & GUARD:
            return x, None, Jr
    else:
        x, Jp = O.filtered(A, B, C, D, P, version, eps_constant, sem=O.SEM_GPU,
                           cyclic=bool(cyclic), rho="auto", return_J=True)
        if not options & GUARD:
            return x, Jp, None
        Jsys = np.repeat(np.asarray(Jp, dtype=np.int32)[:, None], n_sys, axis=1)
    x, rec = O.filtered_guard(A, B, C, D, P, Jsys, eps_constant, cyclic=bool(cyclic))
    return x, Jp, rec | (fb << 16)


def test_cpp_backend_protocol_and_first_solve_graph(tmp_path):
    if not os.path.exists(PROG):
        pytest.fail(f"{PROG} is not built (make -C tests/cpp)")
    inputs = []
    blob = [struct.pack("<i", len(CASES))]
    for k, (kind, n_sys, n_row, P, cyclic, epsc, opt) in enumerate(CASES):
        A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, 0.5, 20.0, seed=40 + k)
        inputs.append((A, B, C, D))
        blob.append(struct.pack("<6id", int(kind), n_sys, n_row, P, cyclic, opt, epsc))
        for x in (A, B, C, D):
            blob.append(np.ascontiguousarray(x, dtype="<f8").tobytes())
    cin, cout = tmp_path / "cases.bin", tmp_path / "results.bin"
    cin.write_bytes(b"".join(blob))
    print(f"[gpu-case] backend_check: {len(CASES)} cases", flush=True)
    sys.stdout.flush()
    res = subprocess.run([PROG, str(cin), str(cout)], capture_output=True, text=True, timeout=100)
    print(res.stdout)
    assert res.returncode == 0, res.stderr
    buf = cout.read_bytes()
    pos = 0
    for (kind, n_sys, n_row, P, cyclic, epsc, opt), (A, B, C, D) in zip(CASES, inputs):
        J = np.frombuffer(buf, "<i4", P, pos)
        pos += 4 * P
        gpu_ms, gpu_ms_off, comm_ms = np.frombuffer(buf, "<f8", 3, pos)
        pos += 24
        parse_ok, graph_equal = np.frombuffer(buf, "<i4", 2, pos)
        pos += 8
        x = np.frombuffer(buf, "<f8", n_sys * n_row, pos).reshape(n_row, n_sys)
        pos += 8 * n_sys * n_row
        rec = None
        if opt:
            rec = np.frombuffer(buf, "<i4", P * n_sys, pos).reshape(P, n_sys)
            pos += 4 * P * n_sys
        ref, Jref, recref = oracle_solution(kind, A, B, C, D, P, cyclic, epsc, opt)
        what = f"{kind.name} n_sys={n_sys} n_row={n_row} P={P} cyclic={cyclic} options={opt}"
        np.testing.assert_array_equal(x, ref, err_msg=what)
        if Jref is not None:
            assert list(J) == list(Jref), what
        if recref is not None:
            np.testing.assert_array_equal(rec, recref, err_msg=what)
        assert gpu_ms > 0.0, what
        assert gpu_ms_off == 0.0, what
        assert comm_ms == 0.0, what
        assert parse_ok == 1
        assert graph_equal == 1, what
    assert pos == len(buf)
