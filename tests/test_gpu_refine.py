"""FP32 elimination with FP64 refinement (ftd_solve_f32_refine) against the
oracle's restatement (ftdo_pascal_ir): x bit for bit after 0..3 sweeps, on
shapes with partial tiles, uneven partitions and every fused kernel shape;
and what the sweeps are for: the FP64 solve of the stored system within
1e-12 after two of them, on the mixed-precision workload's coefficients
(r ~ U[5, 500], rho to 0.499)."""
import sys

import numpy as np
import pytest

import oracle as O
from filtered_tdma_amd import FtdError, TdmaBackendGPU, cases

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def path_Q(path):
    return int(path.split("Q=")[1].split()[0])


def relerr(x, ref):
    num = np.max(np.abs(x - ref), axis=0)
    den = np.maximum(np.max(np.abs(ref), axis=0), 1e-300)
    return float(np.max(num / den))


def run_refine(A32, B32, C32, D32, P, sweeps):
    n_row, n_sys = A32.shape
    be = TdmaBackendGPU("pascal", n_sys, n_row, P)
    try:
        tA, tB, tC, tD = (torch.from_numpy(np.ascontiguousarray(x)).cuda()
                          for x in (A32, B32, C32, D32))
        tX = torch.full((n_row, n_sys), float("nan"), dtype=torch.float64, device="cuda")
        be.solve_f32_refine(tA, tB, tC, tD, tX, sweeps)
        torch.cuda.synchronize()
        return tX.cpu().numpy(), be.last_path()
    finally:
        be.close()


def stored(n_sys, n_row, lo, hi, seed):
    A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, lo, hi, seed=seed)
    return tuple(x.astype(np.float32) for x in (A, B, C, D))


@pytest.mark.parametrize("sweeps", [0, 1, 2, 3])
@pytest.mark.parametrize("n_sys,n_row,P", [(300, 512, 8), (257, 129, 1), (64, 96, 3),
                                           (1024, 512, 8), (2048, 256, 4), (512, 200, 8),
                                           (100, 40, 1), (96, 1000, 1)])
def test_refine_bitwise_vs_oracle(n_sys, n_row, P, sweeps):
    note(f"refine n_sys={n_sys} n_row={n_row} P={P} sweeps={sweeps}")
    A32, B32, C32, D32 = stored(n_sys, n_row, 5.0, 500.0, seed=7 + P)
    x, path = run_refine(A32, B32, C32, D32, P, sweeps)
    assert path.startswith("pascal_ir") and f"sweeps={sweeps}" in path, path
    ref = O.pascal_ir(A32, B32, C32, D32, path_Q(path), sweeps)
    np.testing.assert_array_equal(x, ref, err_msg=path)


def test_refine_reaches_the_fp64_solve_of_the_stored_system():
    """Each sweep shrinks the error by about kappa * 2^-24; two sweeps are
    within 1e-12 of the FP64 solve of the stored system, while the FP32
    solve alone is not (so the elimination really runs in FP32)."""
    n_sys, n_row, P = 4096, 512, 8
    note("refine convergence 4096 x 512 P=8")
    A32, B32, C32, D32 = stored(n_sys, n_row, 5.0, 500.0, seed=0)
    exact = O.thomas(*(x.astype(np.float64) for x in (A32, B32, C32, D32)))
    errs = []
    for sweeps in range(4):
        x, _ = run_refine(A32, B32, C32, D32, P, sweeps)
        errs.append(relerr(x, exact))
    note(f"errors by sweeps: {errs}")
    assert 1e-9 < errs[0] < 1e-4, errs
    assert errs[1] < errs[0] * 1e-3, errs
    assert errs[2] <= 1e-12, errs
    assert errs[3] <= 1e-12, errs


def test_refine_full_size_sampled_columns():
    """The mixed-precision workload at its full 262144 x 512, P = 8 shape,
    two sweeps: 2048 sampled systems bit for bit against the oracle (systems
    are independent, so the oracle runs on the sampled columns only), and
    within 1e-12 of the FP64 Thomas solve of the stored data."""
    n_sys, n_row, P = 262144, 512, 8
    note("refine full size, two sweeps")
    A32, B32, C32, D32 = stored(n_sys, n_row, 5.0, 500.0, seed=0)
    x, path = run_refine(A32, B32, C32, D32, P, 2)
    cols = np.unique(np.concatenate([np.arange(0, n_sys, 131), np.arange(n_sys - 32, n_sys)]))
    x = np.ascontiguousarray(x[:, cols])
    sub = [np.ascontiguousarray(a[:, cols]) for a in (A32, B32, C32, D32)]
    ref = O.pascal_ir(*sub, path_Q(path), 2)
    np.testing.assert_array_equal(x, ref, err_msg=path)
    assert relerr(x, O.thomas(*(a.astype(np.float64) for a in sub))) <= 1e-12


def test_refine_errors():
    A32, B32, C32, D32 = stored(64, 96, 5.0, 500.0, seed=1)
    tA, tB, tC, tD = (torch.from_numpy(x).cuda() for x in (A32, B32, C32, D32))
    tX = torch.empty((96, 64), dtype=torch.float64, device="cuda")
    be = TdmaBackendGPU("pascal", 64, 96, 2)
    try:
        for bad in (-1, 5):
            with pytest.raises(FtdError):
                be.solve_f32_refine(tA, tB, tC, tD, tX, bad)
    finally:
        be.close()
    be = TdmaBackendGPU("filtered", 64, 96, 2)
    try:
        with pytest.raises(FtdError):
            be.solve_f32_refine(tA, tB, tC, tD, tX, 1)
    finally:
        be.close()
