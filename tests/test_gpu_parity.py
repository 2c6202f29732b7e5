"""Parity of the HIP engine (libftdma.so, through its C ABI) with the oracle.

Bar: the engine and the oracle follow the same statements with the same
rounding (both built with -ffp-contract=off), so every comparison below is
bit-for-bit (`assert_array_equal`) where the engine and the oracle use the
same partitioning. Exact solves are additionally checked against serial
Thomas within 1e-13 relative (the north-star bound for PaScaL).
"""
import sys

import numpy as np
import pytest

import oracle as O
from filtered_tdma_amd import FtdError, Kind, TdmaBackendGPU
from filtered_tdma_amd import cases

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()


def run(kind, A, B, C, D, P=1, cyclic=False, rho=False, eps=None, eps_constant=0.0,
        tuning=None, nr=False):
    """Solve on the GPU; returns (x, J list, path, B after)."""
    n_row, n_sys = D.shape
    be = TdmaBackendGPU(kind, n_sys, n_row, P, eps_constant)
    try:
        if tuning is not None:
            be.set_tuning(*tuning)
        tA, tB, tC, tD = dev(A), dev(B), dev(C), dev(D)
        if eps is not None:
            be.set_eps(eps)
        if rho:
            be.set_rho_device(tA, tB, tC)
        if cyclic:
            be.solve_cyclic(tA, tB, tC, tD)
        elif nr:
            be.solve_nr(tA, tB, tC, tD, n_sys, n_row)
        else:
            be.solve(tA, tB, tC, tD)
        torch.cuda.synchronize()
        J = be.last_J() if Kind(be.kind) != Kind.PASCAL and P > 1 else None
        return tD.cpu().numpy(), J, be.last_path(), tB.cpu().numpy()
    finally:
        be.close()


def path_Q(path):
    """partition count of a PaScaL path string ("... Q=8" or "... Q=8 R=16")"""
    return int(path.split("Q=")[1].split()[0])


def relerr(x, ref):
    """max over systems of ||x - ref||_inf / ||ref||_inf"""
    num = np.max(np.abs(x - ref), axis=0)
    den = np.maximum(np.max(np.abs(ref), axis=0), 1e-300)
    return float(np.max(num / den))


# ----------------------------------------------------------------- exact path

PASCAL_SHAPES = [
    # (n_sys, n_row, forced (M, T, S, Q) or None)
    (1000, 512, None),
    (777, 129, None),
    (64, 3, None),
    (300, 100, None),
    (5, 37, None),
    (1, 17, None),
    (300, 1000, None),  # above 512 rows: 32-row partitions in 512-thread groups
    (64, 600, None),
    (513, 96, (16, 512, 16, 0)),
    (200, 240, (32, 512, 16, 0)),
    (130, 500, (64, 256, 32, 0)),
    (99, 60, (8, 512, 64, 0)),
]


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("n_sys,n_row,tune", PASCAL_SHAPES)
def test_pascal_fused_bitwise(n_sys, n_row, tune, cyclic):
    if cyclic and n_row < 3:
        pytest.skip("a periodic system needs at least 3 rows")
    note(f"pascal_fused n_sys={n_sys} n_row={n_row} tune={tune} cyclic={cyclic}")
    A, B, C, D = cases.general_batch(n_sys, n_row, seed=11)
    x, _, path, Bafter = run(Kind.PASCAL, A, B, C, D, P=1, cyclic=cyclic, tuning=tune)
    if path.startswith("thomas"):
        # a system that fits one lane: plain or periodic Thomas, the
        # reference's single-rank path
        assert n_row <= 64
        ref = O.thomas(A, B, C, D, sem=O.SEM_GPU, cyclic=cyclic)
    else:
        assert path.startswith(("pascal_fused", "pascal_stream")), path
        ref = O.pascal(A, B, C, D, path_Q(path), sem=O.SEM_GPU, cyclic=cyclic)
    np.testing.assert_array_equal(x, ref)
    np.testing.assert_array_equal(Bafter, B)
    thomas = O.thomas(A, B, C, D, sem=O.SEM_GPU, cyclic=cyclic)
    assert relerr(x, thomas) <= 1e-13


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("n_sys,n_row,P", [(300, 64, 1), (257, 50, 1), (100, 90, 3), (64, 40, 4)])
def test_exact_general_bitwise(n_sys, n_row, P, cyclic):
    note(f"exact_general n_sys={n_sys} n_row={n_row} P={P} cyclic={cyclic}")
    A, B, C, D = cases.general_batch(n_sys, n_row, seed=5)
    x, _, path, _ = run(Kind.PASCAL, A, B, C, D, P=P, cyclic=cyclic, tuning=(-1, 0, 0, 0))
    assert "general" in path, path
    if P == 1:
        ref = O.thomas(A, B, C, D, sem=O.SEM_GPU, cyclic=cyclic)
    else:
        ref = O.pascal(A, B, C, D, P, sem=O.SEM_GPU, cyclic=cyclic)
    np.testing.assert_array_equal(x, ref)


def test_thomas_tiny_rows():
    note("thomas n_row in {1, 2}")
    for n_row in (1, 2):
        A, B, C, D = cases.general_batch(33, n_row, seed=2)
        x, _, path, _ = run(Kind.PASCAL, A, B, C, D)
        assert path == "thomas_general"
        np.testing.assert_array_equal(x, O.thomas(A, B, C, D, sem=O.SEM_GPU))


def test_solve_nr_other_dims():
    note("solve_nr with dims different from the handle")
    A, B, C, D = cases.general_batch(70, 45, seed=8)
    be = TdmaBackendGPU("pascal", 10, 10, 1)
    try:
        tA, tB, tC, tD = dev(A), dev(B), dev(C), dev(D)
        be.solve_nr(tA, tB, tC, tD, 70, 45)
        torch.cuda.synchronize()
        path = be.last_path()
        if path.startswith("thomas_fused"):
            ref = O.thomas(A, B, C, D, sem=O.SEM_GPU)
        else:
            ref = O.pascal(A, B, C, D, path_Q(path), sem=O.SEM_GPU)
        np.testing.assert_array_equal(tD.cpu().numpy(), ref)
    finally:
        be.close()


# ------------------------------------------- LDS-streamed persistent kernels

STREAM_SHAPES = [
    # (n_sys, n_row, P): n_sys % 32 == 0; variants (M, Q) = (64, 8), (32, 8),
    # (64, 4), (32, 4), (32, 16); several tiles per resident block in the
    # first case, fewer tiles than blocks in others, uneven para_range splits
    (16480, 509, 8),   # M=64, 515 tiles, m_p = 64/63
    (96, 250, 8),      # M=32, 3 tiles, m_p = 32/31
    (24, 24, 8),       # M=32 (m_p = 3, rows past m_p stream clamped), n_sys < 32 -> fused
    (4128, 256, 4),    # M=64, Q=4
    (2048, 100, 4),    # M=32, Q=4, m_p = 25
    (3200, 500, 16),   # M=32, Q=16, m_p = 32/31
    (8192, 512, 1),    # P=1: 8 internal partitions of 64 rows
    (9216, 24, 8),     # M=32 with m_p = 3
]


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("n_sys,n_row,P", STREAM_SHAPES)
def test_pascal_stream_bitwise(n_sys, n_row, P, cyclic):
    """The streamed kernel equals the oracle's PaScaL of the same partitioning
    bit for bit, and so equals the register-resident fused kernel."""
    note(f"pascal_stream n_sys={n_sys} n_row={n_row} P={P} cyclic={cyclic}")
    A, B, C, D = cases.general_batch(n_sys, n_row, seed=21)
    x, _, path, Bafter = run(Kind.PASCAL, A, B, C, D, P=P, cyclic=cyclic)
    if n_sys % 32 == 0:
        assert path.startswith("pascal_stream"), path
    Q = path_Q(path)
    assert Q == (P if P > 1 else 8)
    np.testing.assert_array_equal(x, O.pascal(A, B, C, D, Q, sem=O.SEM_GPU, cyclic=cyclic))
    np.testing.assert_array_equal(Bafter, B)
    xf, _, pathf, _ = run(Kind.PASCAL, A, B, C, D, P=P, cyclic=cyclic, tuning=(-2, 0, 0, 0))
    assert pathf.startswith("pascal_fused"), pathf
    np.testing.assert_array_equal(x, xf)


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("kind", [Kind.FILTERED, Kind.FILTERED_V2])
@pytest.mark.parametrize("n_sys,n_row,P,rr", [
    (16480, 509, 8, (0.5, 20.0)),
    (8448, 512, 8, (0.05, 0.3)),   # interior J
    (96, 250, 8, (0.1, 1.0)),
    (4128, 256, 4, (0.05, 0.3)),
    (2048, 100, 4, (0.5, 20.0)),
    (3200, 500, 16, (0.05, 0.3)),
    (9216, 24, 8, (0.1, 1.0)),
])
def test_filtered_stream_bitwise(n_sys, n_row, P, rr, kind, cyclic):
    version = 1 if kind == Kind.FILTERED else 2
    A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, rr[0], rr[1], seed=17)
    for eps_c in (1e-3, 1e-12):
        note(f"filtered_stream v{version} n_sys={n_sys} n_row={n_row} P={P} r={rr} "
             f"cyclic={cyclic} eps_constant={eps_c}")
        x, J, path, Bafter = run(kind, A, B, C, D, P=P, cyclic=cyclic, rho=True,
                                 eps_constant=eps_c)
        assert path.startswith(f"filtered_v{version}_stream"), path
        ref, Jref = O.filtered(A, B, C, D, P, version, eps_c, sem=O.SEM_GPU, cyclic=cyclic,
                               rho="auto", return_J=True)
        assert list(J) == list(Jref)
        np.testing.assert_array_equal(x, ref)
        np.testing.assert_array_equal(Bafter, B)
        xf, Jf, pathf, _ = run(kind, A, B, C, D, P=P, cyclic=cyclic, rho=True,
                               eps_constant=eps_c, tuning=(-2, 0, 0, 0))
        assert "fused" in pathf, pathf
        assert list(Jf) == list(J)
        np.testing.assert_array_equal(x, xf)


def test_stream_J_edges():
    """J = 0 (tiny RHS), interior J and J = m-1 through the streamed kernels."""
    n_sys, n_row, P = 4096, 256, 8
    A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, 0.05, 0.3, seed=19)
    for scale, eps in ((1e-14, 1e-10), (1.0, 1e-10), (1.0, 1e-300)):
        for kind, version in ((Kind.FILTERED_V2, 2), (Kind.FILTERED, 1)):
            note(f"filtered_stream v{version} D scale={scale} eps={eps}")
            Ds = D * scale
            x, J, path, _ = run(kind, A, B, C, Ds, P=P, rho=True, eps=eps)
            assert "stream" in path, path
            ref, Jref = O.filtered(A, B, C, Ds, P, version, eps, eps_direct=True,
                                   sem=O.SEM_GPU, rho="auto", return_J=True)
            assert list(J) == list(Jref)
            np.testing.assert_array_equal(x, ref)


def test_stream_needs_aligned_arrays():
    """An array that is not 16-byte aligned takes the fused kernel, same bits."""
    n_sys, n_row, P = 2048, 256, 8
    note("misaligned D -> fused")
    A, B, C, D = cases.general_batch(n_sys, n_row, seed=23)
    be = TdmaBackendGPU("pascal", n_sys, n_row, P)
    try:
        tA, tB, tC = dev(A), dev(B), dev(C)
        flat = torch.zeros(n_row * n_sys + 1, dtype=torch.float64, device="cuda")
        tD = flat[1:].view(n_row, n_sys)
        tD.copy_(dev(D))
        be.solve(tA, tB, tC, tD)
        torch.cuda.synchronize()
        assert be.last_path().startswith("pascal_fused"), be.last_path()
        np.testing.assert_array_equal(tD.cpu().numpy(), O.pascal(A, B, C, D, P, sem=O.SEM_GPU))
    finally:
        be.close()


# --------------------------------------------------------------- filtered path

FILTERED_CASES = [
    # (n_sys, n_row, P, r range) -- m = n_row / P rows per partition
    (300, 64, 8, (0.5, 20.0)),     # m = 8
    (257, 128, 8, (0.05, 0.3)),    # m = 16, interior J
    (200, 256, 8, (0.5, 20.0)),    # m = 32
    (130, 512, 8, (0.5, 20.0)),    # m = 64 (the 262144 x 512 workload's partitions)
    (96, 100, 3, (0.05, 0.3)),     # uneven para_range split, m = 34/33
    (70, 40, 2, (0.1, 1.0)),
]


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("kind", [Kind.FILTERED, Kind.FILTERED_V2])
@pytest.mark.parametrize("n_sys,n_row,P,rr", FILTERED_CASES)
def test_filtered_fused_bitwise(n_sys, n_row, P, rr, kind, cyclic):
    version = 1 if kind == Kind.FILTERED else 2
    for eps_c, rho in ((1e-3, True), (1e-12, True), (1e-6, False)):
        note(f"filtered v{version} n_sys={n_sys} n_row={n_row} P={P} r={rr} cyclic={cyclic} "
             f"eps_constant={eps_c} rho={rho}")
        A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, rr[0], rr[1], seed=7)
        x, J, path, Bafter = run(kind, A, B, C, D, P=P, cyclic=cyclic, rho=rho,
                                 eps_constant=eps_c)
        assert "fused" in path or "stream" in path, path
        ref, Jref = O.filtered(A, B, C, D, P, version, eps_c, sem=O.SEM_GPU, cyclic=cyclic,
                               rho="auto" if rho else None, return_J=True)
        assert list(J) == list(Jref)
        np.testing.assert_array_equal(x, ref)
        np.testing.assert_array_equal(Bafter, B)


@pytest.mark.parametrize("kind", [Kind.FILTERED, Kind.FILTERED_V2])
@pytest.mark.parametrize("n_sys,n_row,P", [(100, 300, 2), (64, 96, 3), (33, 160, 2)])
def test_filtered_general_bitwise(n_sys, n_row, P, kind):
    version = 1 if kind == Kind.FILTERED else 2
    for cyclic in (False, True):
        note(f"filtered_general v{version} n_sys={n_sys} n_row={n_row} P={P} cyclic={cyclic}")
        A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, 0.05, 0.5, seed=3)
        x, J, path, _ = run(kind, A, B, C, D, P=P, cyclic=cyclic, rho=True, eps_constant=1e-4,
                            tuning=(-1, 0, 0, 0))
        assert "general" in path, path
        ref, Jref = O.filtered(A, B, C, D, P, version, 1e-4, sem=O.SEM_GPU, cyclic=cyclic,
                               rho="auto", return_J=True)
        assert list(J) == list(Jref)
        np.testing.assert_array_equal(x, ref)


def test_filtered_direct_eps_and_J_edges():
    """eps_ given directly; J = 0 (tiny RHS), interior J and J = m-1."""
    n_sys, n_row, P = 128, 128, 4
    A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, 0.05, 0.3, seed=9)
    for scale, eps in ((1e-14, 1e-10), (1.0, 1e-10), (1.0, 1e-300)):
        for kind, version in ((Kind.FILTERED_V2, 2), (Kind.FILTERED, 1)):
            note(f"filtered v{version} D scale={scale} eps={eps}")
            Ds = D * scale
            x, J, _, _ = run(kind, A, B, C, Ds, P=P, rho=True, eps=eps)
            ref, Jref = O.filtered(A, B, C, Ds, P, version, eps, eps_direct=True,
                                   sem=O.SEM_GPU, rho="auto", return_J=True)
            assert list(J) == list(Jref)
            np.testing.assert_array_equal(x, ref)


def test_filtered_zero_rhs():
    note("filtered v2 with an all-zero RHS (J defined as 0)")
    A, B, C, D, _ = cases.toeplitz_batch(64, 64, 0.1, 0.4, seed=1)
    D = np.zeros_like(D)
    x, J, _, _ = run(Kind.FILTERED_V2, A, B, C, D, P=4, rho=True, eps=1e-10)
    assert J == [0, 0, 0, 0]
    assert np.all(x == 0.0)


def test_filtered_P1_is_thomas():
    note("filtered with one partition -> Thomas")
    A, B, C, D, _ = cases.toeplitz_batch(100, 70, 0.5, 5.0, seed=4)
    for kind in (Kind.FILTERED, Kind.FILTERED_V2):
        x, _, path, _ = run(kind, A, B, C, D, P=1, rho=True, eps_constant=1e-3)
        assert relerr(x, O.thomas(A, B, C, D)) <= 1e-13, path


# ------------------------------------------------------------------ errors

def test_errors_are_raised():
    note("error statuses")
    A, B, C, D = cases.general_batch(16, 8, seed=1)
    with pytest.raises(FtdError):
        run(Kind.FILTERED_V2, A, B, C, D, P=4)  # 2 rows per partition
    with pytest.raises(FtdError):
        TdmaBackendGPU("pascal", 4, 4, 8)  # more partitions than rows
    be = TdmaBackendGPU("pascal", 16, 8, 1)
    try:
        with pytest.raises(FtdError):
            be.solve(0, 0, 0, 0)
    finally:
        be.close()


# ------------------------------------------- full 262144 x 512 workload shape

@pytest.mark.parametrize("kind", [Kind.PASCAL, Kind.FILTERED_V2, Kind.FILTERED])
def test_full_size_columns_match_oracle(kind):
    """At 262144 x 512 (P = 8) every system is independent, so a sample of
    columns solved by the oracle must equal the same columns of the GPU solve
    bit for bit. For the filtered kinds J is a max over all systems: the
    oracle takes the GPU's J for the sampled columns, and J itself is checked
    against the formula applied to the full-size maxima."""
    n_sys, n_row, P, eps = 262144, 512, 8, 1e-10
    note(f"full-size {kind.name} n_sys={n_sys} n_row={n_row} P={P}")
    A, B, C, D, r = cases.toeplitz_batch(n_sys, n_row, 0.5, 20.0, seed=0)
    x, J, path, _ = run(kind, A, B, C, D, P=P, rho=True, eps=eps)
    cols = np.unique(np.concatenate([np.arange(0, n_sys, 4099), [0, n_sys - 1]]))
    sub = [np.ascontiguousarray(a[:, cols]) for a in (A, B, C, D)]
    if kind == Kind.PASCAL:
        ref = O.pascal(*sub, path_Q(path), sem=O.SEM_GPU)
        np.testing.assert_array_equal(x[:, cols], ref)
        # residual of the whole batch
        res = B * x - D
        res[1:] += A[1:] * x[:-1]
        res[:-1] += C[:-1] * x[1:]
        assert float(np.max(np.abs(res))) < 1e-12
    else:
        version = 1 if kind == Kind.FILTERED else 2
        rho = O.set_rho(A, B, C)
        ref = O.filtered(*sub, P, version, eps, eps_direct=True, sem=O.SEM_GPU,
                         rho=rho, J_force=J)
        np.testing.assert_array_equal(x[:, cols], ref)
        if version == 2:
            for p in range(P):
                s, e = O.para_range(0, n_row - 1, P, p)
                mb = max(np.max(np.abs(D[s])), np.max(np.abs(D[e])))
                rh = max(np.max(rho[0][s + 1:e + 1]), np.max(rho[1][s + 1:e + 1]))
                assert J[p] == O.cal_J(rh, mb, eps, e - s + 1)


# ------------------------------------------------------- FP32 storage (C3)

@pytest.mark.parametrize("d32", [False, True])
@pytest.mark.parametrize("n_sys,n_row,P", [(300, 512, 8), (257, 129, 1), (64, 96, 3),
                                           (1024, 512, 8), (2048, 256, 4), (512, 200, 8),
                                           (2048, 512, 16)])
def test_f32_storage_bitwise_vs_oracle_on_widened_data(n_sys, n_row, P, d32):
    """FP32 A, B, C (and D): the engine eliminates in FP64 on the exactly
    widened values, so the result equals the FP64 PaScaL of the stored data."""
    note(f"f32 storage n_sys={n_sys} n_row={n_row} P={P} d32={d32}")
    A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, 5.0, 500.0, seed=12)
    A32, B32, C32 = (x.astype(np.float32) for x in (A, B, C))
    D32 = D.astype(np.float32)
    be = TdmaBackendGPU("pascal", n_sys, n_row, P)
    try:
        tA, tB, tC = (torch.from_numpy(x).cuda() for x in (A32, B32, C32))
        if d32:
            tD = torch.from_numpy(D32).cuda()
            tX = torch.empty((n_row, n_sys), dtype=torch.float64, device="cuda")
            be.solve_f32(tA, tB, tC, tD, tX)
        else:
            tX = dev(D)
            be.solve_f32c(tA, tB, tC, tX)
        torch.cuda.synchronize()
        path = be.last_path()
        x = tX.cpu().numpy()
    finally:
        be.close()
    if d32 and n_sys % 32 == 0 and P in (4, 8, 16):
        assert path.startswith("pascal_stream_f32"), path  # the streamed kernel, same bits
    Q = path_Q(path)
    Dw = D32.astype(np.float64) if d32 else D
    ref = O.pascal(A32.astype(np.float64), B32.astype(np.float64), C32.astype(np.float64), Dw, Q,
                   sem=O.SEM_GPU)
    np.testing.assert_array_equal(x, ref)
    assert relerr(x, O.thomas(A32.astype(np.float64), B32.astype(np.float64),
                              C32.astype(np.float64), Dw)) <= 1e-12


@pytest.mark.parametrize("d32", [False, True])
def test_f32_storage_full_size_sampled_columns(d32):
    """FP32 storage at the full 262144 x 512, P = 8 shape with r ~ U[5, 500]
    (the streamed FP32 kernel walks many tiles per workgroup): 2048 sampled
    systems bit for bit against the oracle's FP64 PaScaL of the widened
    stored data, and within 1e-12 of its serial Thomas."""
    n_sys, n_row, P = 262144, 512, 8
    note(f"f32 storage full size d32={d32}")
    A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, 5.0, 500.0, seed=0)
    A32, B32, C32 = (x.astype(np.float32) for x in (A, B, C))
    D32 = D.astype(np.float32) if d32 else None
    del A, B, C
    be = TdmaBackendGPU("pascal", n_sys, n_row, P)
    try:
        tA, tB, tC = (torch.from_numpy(x).cuda() for x in (A32, B32, C32))
        if d32:
            tD = torch.from_numpy(D32).cuda()
            tX = torch.empty((n_row, n_sys), dtype=torch.float64, device="cuda")
            be.solve_f32(tA, tB, tC, tD, tX)
        else:
            tX = dev(D)
            be.solve_f32c(tA, tB, tC, tX)
        torch.cuda.synchronize()
        path = be.last_path()
        cols = np.unique(np.concatenate([np.arange(0, n_sys, 131), np.arange(n_sys - 32, n_sys)]))
        x = tX.cpu().numpy()[:, cols]
        del tA, tB, tC, tX
    finally:
        be.close()
    if d32:
        assert path.startswith("pascal_stream_f32"), path
    sub = [np.ascontiguousarray(a[:, cols]).astype(np.float64) for a in (A32, B32, C32)]
    Dw = np.ascontiguousarray((D32 if d32 else D)[:, cols]).astype(np.float64)
    ref = O.pascal(*sub, Dw, path_Q(path), sem=O.SEM_GPU)
    np.testing.assert_array_equal(x, ref)
    assert relerr(x, O.thomas(*sub, Dw)) <= 1e-12
