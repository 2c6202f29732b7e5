"""Learned-J filtered mode (ftd_set_j_mode(h, 1)) against the oracle's
restatement (ftdo_filtered_learned): x and the J of every (partition,
system), fallback flag included, bit for bit; and the accuracy it buys
against the reference's cal_J solve of the same systems."""
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


def mixed_scale(n_sys, n_row, seed, r=(0.05, 3.0)):
    """Toeplitz systems whose right-hand sides span 10 decades, so one eps_
    leaves some systems with a short J, some with a long one and some on
    the fallback."""
    A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, r[0], r[1], seed=seed)
    scale = 10.0 ** np.linspace(-8, 2, n_sys)
    return A, B, C, D * scale[None, :]


def run_learned(A, B, C, D, P, eps_c, cyclic=False, complex_rhs=False):
    n_row, n_sys = A.shape
    be = TdmaBackendGPU(Kind.FILTERED, n_sys, n_row, P, eps_c)
    try:
        be.set_j_mode(1)
        tA, tB, tC = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (A, B, C))
        tD = torch.from_numpy(np.ascontiguousarray(D)).cuda()
        if complex_rhs:
            be.solve_complex(tA, tB, tC, tD)
        elif cyclic:
            be.solve_cyclic(tA, tB, tC, tD)
        else:
            be.solve(tA, tB, tC, tD)
        torch.cuda.synchronize()
        J, fb = be.last_J_systems(complex_rhs)
        return tD.cpu().numpy(), J, fb, be.last_path()
    finally:
        be.close()


@pytest.mark.parametrize("n_sys,n_row,P,cyclic", [
    (300, 64, 8, False),     # m = 8
    (257, 128, 8, False),    # m = 16, odd n_sys (dead lanes)
    (512, 256, 8, False),    # m = 32
    (130, 512, 8, False),    # m = 64
    (200, 96, 4, True),      # periodic
    (100, 131, 4, False),    # uneven partitions 33/33/33/32
    (1024, 512, 8, False),   # the streamed pass (m = 64, n_sys % 32 == 0)
    (2048, 200, 4, False),   # streamed, partitions of 50 rows
    (512, 384, 8, True),     # streamed, periodic
    (1024, 509, 8, False),   # streamed, partitions of 64 and 63 rows in one wave
    (200, 300, 2, False),    # partitions of 150 rows: the general kernels
    (130, 1000, 4, True),    # general, periodic, m = 250
    (257, 520, 3, False),    # general, uneven partitions 174/173/173
])
@pytest.mark.parametrize("eps_c", [1e-2, 1e-8])
def test_learned_bitwise(n_sys, n_row, P, cyclic, eps_c):
    note(f"learned n_sys={n_sys} n_row={n_row} P={P} cyclic={cyclic} eps_c={eps_c}")
    A, B, C, D = mixed_scale(n_sys, n_row, seed=n_row + P)
    X, J, fb, path = run_learned(A, B, C, D, P, eps_c, cyclic=cyclic)
    m_max = (n_row + P - 1) // P
    streamed = n_sys % 32 == 0 and 32 < m_max <= 64
    shape = "stream" if streamed else ("fused" if m_max <= 64 else "general")
    assert path.startswith("filtered_learned_" + shape), path
    xr, Jr = O.filtered_learned(A, B, C, D, P, eps_c, cyclic=cyclic)
    np.testing.assert_array_equal(J, Jr & 0xFFFF)
    np.testing.assert_array_equal(fb, Jr >> 16)
    np.testing.assert_array_equal(X, xr)


def test_learned_complex_keeps_a_J_per_part():
    n_sys, n_row, P = 150, 128, 4
    A, B, C, Dr = mixed_scale(n_sys, n_row, seed=5)
    Di = mixed_scale(n_sys, n_row, seed=6)[3][:, ::-1] * 1e-3
    Z = Dr + 1j * np.ascontiguousarray(Di)
    note(f"learned complex n_sys={n_sys} n_row={n_row} P={P}")
    X, J, fb, path = run_learned(A, B, C, Z, P, 1e-6, complex_rhs=True)
    assert path.startswith("filtered_learned_fused"), path
    for par, part, x in ((0, Z.real, X.real), (1, Z.imag, X.imag)):
        xr, Jr = O.filtered_learned(A, B, C, np.ascontiguousarray(part), P, 1e-6)
        np.testing.assert_array_equal(x, xr)
        np.testing.assert_array_equal(J[:, par::2], Jr & 0xFFFF)
        np.testing.assert_array_equal(fb[:, par::2], Jr >> 16)


def test_learned_accuracy_and_work_against_cal_J():
    """Same systems, same eps_: the learned mode's error stays within the
    cal_J solve's error scale, with a smaller mean J."""
    n_sys, n_row, P, eps_c = 2048, 256, 8, 1e-4
    A, B, C, D = mixed_scale(n_sys, n_row, seed=9, r=(0.05, 0.4))
    note(f"learned vs cal_J n_sys={n_sys} n_row={n_row} P={P}")
    X, J, fb, _ = run_learned(A, B, C, D, P, eps_c)
    xe = O.thomas(A, B, C, D)
    be = TdmaBackendGPU(Kind.FILTERED, n_sys, n_row, P, eps_c)
    try:
        tA, tB, tC = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (A, B, C))
        tD = torch.from_numpy(np.ascontiguousarray(D)).cuda()
        be.set_rho_device(tA, tB, tC)
        be.solve(tA, tB, tC, tD)
        torch.cuda.synchronize()
        Xc = tD.cpu().numpy()
        Jc = be.last_J(P)
    finally:
        be.close()
    m = n_row // P
    eps = eps_c / (m * P) ** 2
    err_l = np.max(np.abs(X - xe), axis=0)
    err_c = np.max(np.abs(Xc - xe), axis=0)
    # per system: within a few eps_ of the exact solve, or no worse than cal_J's
    assert np.all((err_l <= 50 * eps) | (err_l <= 2 * err_c + 1e-15)), \
        (err_l.max(), err_c.max(), eps)
    assert J.mean() < np.mean(Jc), (J.mean(), Jc)
