"""Complex right-hand sides (ftd_solve_complex) against the reference's way of
solving them: two real solves, one per part, with the same coefficients
(apps/channel_cpu/PressureSolver.cpp:462-464). Bit for bit against the
oracle's two real solves, per-part J included."""
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


def run_complex(kind, A, B, C, Z, P, rho=False, eps_constant=0.0, eps=None):
    n_row, n_sys = Z.shape
    be = TdmaBackendGPU(kind, n_sys, n_row, P, eps_constant)
    try:
        tA, tB, tC = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (A, B, C))
        tZ = torch.from_numpy(np.ascontiguousarray(Z)).cuda()
        if eps is not None:
            be.set_eps(eps)
        if rho:
            be.set_rho_device(tA, tB, tC)
        be.solve_complex(tA, tB, tC, tZ)
        torch.cuda.synchronize()
        J = be.last_J(2 * P) if Kind(be.kind) != Kind.PASCAL and P > 1 else None
        return tZ.cpu().numpy(), J, be.last_path(), tB.cpu().numpy()
    finally:
        be.close()


def general_complex(n_sys, n_row, seed):
    A, B, C, Dr = cases.general_batch(n_sys, n_row, seed=seed)
    Di = cases.general_batch(n_sys, n_row, seed=seed + 1000)[3]
    return A, B, C, Dr + 1j * Di


@pytest.mark.parametrize("n_sys,n_row,P", [
    (37, 40, 1),      # Thomas in one lane, odd n_sys (dead lanes of both parities)
    (37, 100, 1),     # PaScaL over internal partitions
    (300, 512, 8),
    (1024, 129, 4),   # uneven split 33/32/32/32
    (1024, 256, 1),   # one caller partition: internal partitions of 32 rows
    (2048, 512, 16),  # 16 caller partitions of 32 rows
    (1024, 130, 1),   # internal partitions of 26 rows
])
def test_complex_pascal_equals_two_real_solves(n_sys, n_row, P):
    note(f"complex pascal n_sys={n_sys} n_row={n_row} P={P}")
    A, B, C, Z = general_complex(n_sys, n_row, seed=51)
    X, _, path, Bafter = run_complex(Kind.PASCAL, A, B, C, Z, P)
    if path.startswith("thomas"):
        solve = lambda d: O.thomas(A, B, C, d, sem=O.SEM_GPU)  # noqa: E731
    else:
        solve = lambda d: O.pascal(A, B, C, d, path_Q(path), sem=O.SEM_GPU)  # noqa: E731
    np.testing.assert_array_equal(X.real, solve(np.ascontiguousarray(Z.real)))
    np.testing.assert_array_equal(X.imag, solve(np.ascontiguousarray(Z.imag)))
    np.testing.assert_array_equal(Bafter, B)


@pytest.mark.parametrize("kind", [Kind.FILTERED, Kind.FILTERED_V2])
@pytest.mark.parametrize("n_sys,n_row,P,rr", [
    (301, 256, 8, (0.05, 0.3)),   # interior J, odd n_sys
    (128, 129, 4, (0.5, 20.0)),
])
def test_complex_filtered_keeps_a_J_per_part(kind, n_sys, n_row, P, rr):
    version = 1 if kind == Kind.FILTERED else 2
    A, B, C, Dr, _ = cases.toeplitz_batch(n_sys, n_row, rr[0], rr[1], seed=61)
    Di = cases.toeplitz_batch(n_sys, n_row, rr[0], rr[1], seed=62)[3]
    Di = Di * 1e-3  # a different scale, so the two parts' J differ
    Z = Dr + 1j * Di
    for eps_c in (1e-3, 1e-12):
        note(f"complex filtered v{version} n_sys={n_sys} n_row={n_row} P={P} eps_c={eps_c}")
        X, J, path, _ = run_complex(kind, A, B, C, Z, P, rho=True, eps_constant=eps_c)
        assert "fused" in path, path
        xr, Jr = O.filtered(A, B, C, np.ascontiguousarray(Z.real), P, version, eps_c,
                            sem=O.SEM_GPU, rho="auto", return_J=True)
        xi, Ji = O.filtered(A, B, C, np.ascontiguousarray(Z.imag), P, version, eps_c,
                            sem=O.SEM_GPU, rho="auto", return_J=True)
        assert list(J) == list(Jr) + list(Ji)
        np.testing.assert_array_equal(X.real, xr)
        np.testing.assert_array_equal(X.imag, xi)


def test_complex_channel_poisson_modes():
    """The channel pressure systems (tanh grid, gamma = 2, pinned zero mode)
    at 32 x 32 modes, n_row = 129: pascal (P = 1 and 4) and filtered_v2
    (P = 4, rho from column 0 = the pinned mode, so J is full), each against
    the oracle's two real solves."""
    A, B, C, Z = cases.channel_poisson(32, 32, 129, seed=3)
    for kind, P in ((Kind.PASCAL, 1), (Kind.PASCAL, 4), (Kind.FILTERED_V2, 4)):
        note(f"channel poisson {kind.name} P={P}")
        X, J, path, _ = run_complex(kind, A, B, C, Z, P, rho=True, eps_constant=1e-6)
        Jref = []
        for part, x in ((Z.real, X.real), (Z.imag, X.imag)):
            d = np.ascontiguousarray(part)
            if kind == Kind.PASCAL:
                ref = O.pascal(A, B, C, d, path_Q(path), sem=O.SEM_GPU)
            else:
                ref, Jp = O.filtered(A, B, C, d, P, 2, 1e-6, sem=O.SEM_GPU, rho="auto",
                                     return_J=True)
                Jref += list(Jp)
            np.testing.assert_array_equal(x, ref)
        if J is not None:
            assert J == Jref
            # the last partition holds the upper-wall row of the pinned mode,
            # where |a/b| = 1: its J is full (m_p - 1 = 31) in both parts
            assert J[P - 1] == 31 and J[2 * P - 1] == 31, J
        # the pinned zero mode
        assert X[0, 0] == 0.0


def test_channel_modes_full_size():
    """The channel pressure systems at full size: 128 x 128 modes, n_row =
    129, tanh grid (gamma = 2), pinned zero mode, complex128 RHS ~ N(0, 1).
    pascal (one caller partition), filtered_v2 and learned J (P = 8,
    eps_constant = 1e-6), each part bit for bit against the oracle's real
    solves; learned J's per-system J and fallback flags too."""
    A, B, C, Z = cases.channel_poisson(128, 128, 129, seed=0)
    P, eps_c = 8, 1e-6
    note("channel 128x128 modes pascal P=1")
    X, _, path, _ = run_complex(Kind.PASCAL, A, B, C, Z, 1)
    for part, x in ((Z.real, X.real), (Z.imag, X.imag)):
        d = np.ascontiguousarray(part)
        if path.startswith("thomas"):
            ref = O.thomas(A, B, C, d, sem=O.SEM_GPU)
        else:
            ref = O.pascal(A, B, C, d, path_Q(path), sem=O.SEM_GPU)
        np.testing.assert_array_equal(x, ref)
    assert X[0, 0] == 0.0
    note("channel 128x128 modes filtered_v2 P=8")
    X, J, path, _ = run_complex(Kind.FILTERED_V2, A, B, C, Z, P, rho=True, eps_constant=eps_c)
    Jref = []
    for part, x in ((Z.real, X.real), (Z.imag, X.imag)):
        ref, Jp = O.filtered(A, B, C, np.ascontiguousarray(part), P, 2, eps_c, sem=O.SEM_GPU,
                             rho="auto", return_J=True)
        Jref += list(Jp)
        np.testing.assert_array_equal(x, ref)
    assert J == Jref
    note("channel 128x128 modes learned J P=8")
    n_row, n_sys = Z.shape
    be = TdmaBackendGPU(Kind.FILTERED, n_sys, n_row, P, eps_c)
    try:
        be.set_j_mode(1)
        tA, tB, tC = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (A, B, C))
        tZ = torch.from_numpy(np.ascontiguousarray(Z)).cuda()
        be.solve_complex(tA, tB, tC, tZ)
        torch.cuda.synchronize()
        X = tZ.cpu().numpy()
        Jl, fb = be.last_J_systems(True)
    finally:
        be.close()
    for par, part, x in ((0, Z.real, X.real), (1, Z.imag, X.imag)):
        xr, Jr = O.filtered_learned(A, B, C, np.ascontiguousarray(part), P, eps_c)
        np.testing.assert_array_equal(x, xr)
        np.testing.assert_array_equal(Jl[:, par::2], Jr & 0xFFFF)
        np.testing.assert_array_equal(fb[:, par::2], Jr >> 16)
