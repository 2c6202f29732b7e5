"""The autotuner (ftd_autotune): it times candidate kernel shapes on copies
of D, leaves D untouched, keeps the fastest, and the tuned solve still
equals the oracle bit for bit (same partitioning for P > 1 and the
filtered kinds; the chosen sub-partition count for P == 1)."""
import sys

import numpy as np
import pytest

import oracle as O
from filtered_tdma_amd import Kind, TdmaBackendGPU, cases

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def path_Q(path):
    return int(path.split("Q=")[1].split()[0])


@pytest.mark.parametrize("kind,n_sys,n_row,P", [
    (Kind.PASCAL, 4096, 512, 8),
    (Kind.PASCAL, 3000, 100, 1),
    (Kind.FILTERED_V2, 1000, 128, 4),
    (Kind.FILTERED, 640, 96, 4),
])
def test_autotune_keeps_results(kind, n_sys, n_row, P):
    note(f"autotune {kind.name} n_sys={n_sys} n_row={n_row} P={P}")
    A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, 0.05, 2.0, seed=17)
    be = TdmaBackendGPU(kind, n_sys, n_row, P, 1e-6)
    try:
        tA, tB, tC, tD = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (A, B, C, D))
        be.set_rho_device(tA, tB, tC)
        rep = be.autotune(tA, tB, tC, tD, reps=2)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tD.cpu().numpy(), D)  # D itself is not written
        times = [c["ms"] for c in rep["candidates"]]
        assert len(times) >= 3, rep
        assert rep["best"]["ms"] == min(times)
        assert rep["candidates"][rep["best"]["index"]]["path"] == rep["best"]["path"]
        be.solve(tA, tB, tC, tD)
        torch.cuda.synchronize()
        X = tD.cpu().numpy()
        path = be.last_path()
        assert path == rep["best"]["path"], (path, rep["best"])
        J = be.last_J(P) if kind != Kind.PASCAL else None
    finally:
        be.close()
    if kind == Kind.PASCAL:
        if path.startswith("thomas"):
            ref = O.thomas(A, B, C, D, sem=O.SEM_GPU)
        else:
            Q = path_Q(path)
            assert P == 1 or Q == P
            ref = O.pascal(A, B, C, D, Q, sem=O.SEM_GPU)
        np.testing.assert_array_equal(X, ref)
    else:
        version = 1 if kind == Kind.FILTERED else 2
        ref, Jr = O.filtered(A, B, C, D, P, version, 1e-6, sem=O.SEM_GPU, rho="auto",
                             return_J=True)
        np.testing.assert_array_equal(X, ref)
        assert list(J) == list(Jr)
