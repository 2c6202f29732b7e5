"""Uneven batches (ftd_solve_batched) against the oracle, system by system.

A system of N <= 48 rows is plain Thomas in one lane, a longer one (up to
128 * 48 = 6144 rows) PaScaL with Q = ceil(N/48) partitions inside one
128-lane work item (its 2Q-row reduced system solved by PaScaL over groups
of rows once it has 16 rows or more), a longer one still the lane-per-system
Thomas kernel. Each is compared bit for bit
with the oracle's solve of the same partitioning (GPU semantics), and with
serial Thomas within 1e-13 relative."""
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


def general_uneven(n_rows, seed, gap):
    """Diagonally dominant general coefficients, systems separated by `gap`
    untouched elements (filled with a marker)."""
    n_rows = np.asarray(n_rows, dtype=np.int32)
    off = np.zeros(n_rows.size, dtype=np.int64)
    pos = gap
    for s, n in enumerate(n_rows):
        off[s] = pos
        pos += int(n) + gap
    total = pos
    idx = np.arange(total, dtype=np.uint64)
    A = 2.0 * cases.uniform(seed, 2, idx) - 1.0
    C = 2.0 * cases.uniform(seed, 3, idx) - 1.0
    B = (np.abs(A) + np.abs(C)) * (1.2 + cases.uniform(seed, 4, idx))
    B = np.where(cases.uniform(seed, 5, idx) < 0.5, -B, B)
    D = 2.0 * cases.uniform(seed, 1, idx) - 1.0
    return off, n_rows, A, B, C, D


CHUNK = 48  # rows per lane of the batched kernel (the default)


def red_groups(NR):
    """Groups the kernel solves a reduced system of NR rows in: 1 below 16
    rows, else the largest G with 2 G^2 <= NR."""
    if NR < 16:
        return 1
    G = 1
    while 2 * (G + 1) ** 2 <= NR:
        G += 1
    return G


def aligned_starts(off, n, chunk):
    """The engine's chunk starts for a system of n > chunk rows at element
    offset off: every chunk after the first starts on a 16-element boundary
    of the batch arrays; the first runs to chunk - 16 rows past the first
    boundary (chunk rows where the system starts on one), the others are
    chunk rows, and a last chunk under 16 rows takes 16 from the one before.
    None where the system keeps the para_range split (a remainder under 16
    rows after the first chunk, or more than 128 chunks)."""
    a = (16 - off % 16) % 16
    f = a + chunk - 16 if a else chunk
    rem = n - f
    if rem < 16:
        return None
    Q = 1 + (rem + chunk - 1) // chunk
    if Q > 128:
        return None
    st = [0] + [f + (q - 1) * chunk for q in range(1, Q)] + [n]
    if Q >= 3 and n - st[Q - 1] < 16:
        st[Q - 1] -= 16
    return st


def test_aligned_starts_cut_chunks_on_lines():
    for off in range(0, 40):
        for n in (33, 47, 48, 49, 64, 65, 100, 257, 1000, 4096):
            for chunk in (32, 48):
                if n <= chunk:
                    continue
                st = aligned_starts(off, n, chunk)
                if st is None:
                    assert n < chunk + 16 or n > 127 * chunk, (off, n, chunk)
                    continue
                lens = np.diff(st)
                assert st[0] == 0 and st[-1] == n
                assert lens.min() >= 16 and lens.max() <= chunk, (off, n, chunk, st)
                assert all((off + x) % 16 == 0 for x in st[1:-1]), (off, n, chunk, st)


def oracle_solve(off, n_rows, A, B, C, D, chunk=CHUNK):
    """x of every system by the oracle, with the engine's partitioning:
    chunks cut on 16-element boundaries (aligned_starts) or, where that
    split does not apply, Q = ceil(N / chunk) para_range chunks; the 2Q-row
    reduced system in red_groups(2Q) groups."""
    X = D.copy()
    for s in range(n_rows.size):
        o, n = int(off[s]), int(n_rows[s])
        sl = slice(o, o + n)
        a, b, c, d = (v[sl].reshape(n, 1).copy() for v in (A, B, C, D))
        Q = (n + chunk - 1) // chunk
        st = aligned_starts(o, n, chunk) if 2 <= Q <= 128 else None
        if Q == 1 or Q > 128:
            x = O.thomas(a, b, c, d, sem=O.SEM_GPU)
        elif st is not None:
            Qa = len(st) - 1
            x = O.pascal_starts(a, b, c, d, st, red_groups(2 * Qa), sem=O.SEM_GPU)
        else:
            x = O.pascal_nested(a, b, c, d, Q, red_groups(2 * Q), sem=O.SEM_GPU)
        X[sl] = x[:, 0]
    return X


def thomas_all(off, n_rows, A, B, C, D):
    X = D.copy()
    for s in range(n_rows.size):
        o, n = int(off[s]), int(n_rows[s])
        sl = slice(o, o + n)
        X[sl] = O.thomas(*(v[sl].reshape(n, 1).copy() for v in (A, B, C, D)))[:, 0]
    return X


def run_batched(be, off, n_rows, A, B, C, D):
    tA, tB, tC, tD = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (A, B, C, D))
    be.solve_batched(tA, tB, tC, tD, off, n_rows)
    torch.cuda.synchronize()
    return tA.cpu().numpy(), tB.cpu().numpy(), tD.cpu().numpy(), be.last_path()


EDGE = [1, 2, 3, 4, 16, 31, 32, 33, 34, 47, 48, 49, 63, 64, 65, 66, 95, 96, 97, 127, 128, 129,
        191, 1000, 2048, 2049, 4095, 4096, 4097, 6000, 6143, 6144, 6145, 7000]


def test_batched_uneven_bitwise():
    n_rows = np.concatenate([cases.uneven_lengths(700, 16, 4096, seed=3), EDGE]).astype(np.int32)
    off, n_rows, A, B, C, D = general_uneven(n_rows, seed=31, gap=5)
    note(f"batched {n_rows.size} systems, {int(n_rows.sum())} rows")
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        Aa, Ba, X, path = run_batched(be, off, n_rows, A, B, C, D)
    finally:
        be.close()
    assert path.startswith("batched"), path
    ref = oracle_solve(off, n_rows, A, B, C, D)
    np.testing.assert_array_equal(X, ref)          # gaps untouched, every system bitwise
    np.testing.assert_array_equal(Aa, A)
    np.testing.assert_array_equal(Ba, B)
    th = thomas_all(off, n_rows, A, B, C, D)
    for s in range(n_rows.size):
        sl = slice(int(off[s]), int(off[s]) + int(n_rows[s]))
        err = np.max(np.abs(X[sl] - th[sl])) / max(np.max(np.abs(th[sl])), 1e-300)
        assert err <= 1e-13, (s, int(n_rows[s]), err)


def test_batched_plan_reuse_and_relayout():
    """Same layout twice (cached plan, new values), then a different layout."""
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        for seed, lengths in ((41, [70, 5, 300, 64, 4096]), (42, [70, 5, 300, 64, 4096]),
                              (43, [2000, 129, 33, 1, 1, 65])):
            note(f"batched layout {lengths} seed {seed}")
            off, n_rows, A, B, C, D = general_uneven(lengths, seed=seed, gap=1)
            _, _, X, path = run_batched(be, off, n_rows, A, B, C, D)
            np.testing.assert_array_equal(X, oracle_solve(off, n_rows, A, B, C, D))
    finally:
        be.close()


def test_batched_many_items_per_workgroup():
    """23000 systems (1.5e7 rows): several thousand work items, so every
    workgroup of the resident grid runs a sequence of them, each item's
    leading rows requested while the previous one finishes, across the
    switch from chunked to single-chunk items."""
    n_rows = np.concatenate([cases.uneven_lengths(20000, 16, 4096, seed=1),
                             np.arange(1, 16).repeat(200)]).astype(np.int32)
    np.random.default_rng(2).shuffle(n_rows)
    off, n_rows, A, B, C, D = general_uneven(n_rows, seed=31, gap=1)
    note(f"batched {n_rows.size} systems, {int(n_rows.sum())} rows")
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        _, _, X, path = run_batched(be, off, n_rows, A, B, C, D)
    finally:
        be.close()
    assert path.startswith("batched"), path
    np.testing.assert_array_equal(X, oracle_solve(off, n_rows, A, B, C, D))


@pytest.mark.parametrize("lengths,gap", [
    ([40, 7, 1], 1),      # a 1-row system last
    ([300, 31], 1),       # an odd single-chunk system last
    ([70, 5, 2049], 1),   # chunked, chunks of 48 and 47 rows, the last odd
    ([99], 0),            # one chunked system, three chunks of 33 rows
    ([1], 1),             # one 1-row system at offset 1
    ([1], 0),             # one 1-row system at offset 0 (lane-per-system kernel)
    ([2, 1], 0),          # two short systems back to back, 1-row last
])
def test_batched_arrays_end_at_last_row(lengths, gap):
    """The arrays end with the last system's last row: the staging reads two
    rows at a time and must not run past it, nor drop that row."""
    off, n_rows, A, B, C, D = general_uneven(lengths, seed=len(lengths) + 50, gap=gap)
    end = int(off[-1]) + int(n_rows[-1])
    A, B, C, D = (np.ascontiguousarray(v[:end]) for v in (A, B, C, D))
    note(f"batched array end {lengths}")
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        _, _, X, _ = run_batched(be, off, n_rows, A, B, C, D)
    finally:
        be.close()
    np.testing.assert_array_equal(X, oracle_solve(off, n_rows, A, B, C, D))


def test_batched_reduced_denominator_overflow():
    """A reduced-system denominator that overflows to -inf. System 0 (200
    rows, 5 partitions of 40 rows) has a zero right-hand side, c = 1e200
    on partition 0's last row and a = 1e200, c = 0 on partition 1's first:
    the reduced rows' product overflows, IEEE division gives inv = -0 and x
    stays exactly zero. A reciprocal that is not the division's (v_rcp_f64
    and Newton steps alone give NaN for -inf) would show here."""
    off, n_rows, A, B, C, D = general_uneven([200, 300, 40, 5], seed=77, gap=1)
    o = int(off[0])
    D[o:o + 200] = 0.0
    C[o + 39], A[o + 40], C[o + 40] = 1e200, 1e200, 0.0
    note("batched reduced-solve division fallback")
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        _, _, X, _ = run_batched(be, off, n_rows, A, B, C, D)
    finally:
        be.close()
    ref = oracle_solve(off, n_rows, A, B, C, D)
    assert np.all(ref[o:o + 200] == 0.0)
    np.testing.assert_array_equal(X, ref)


def run_batched_f32(be, off, n_rows, A, B, C, D):
    """FP32 storage; returns x (float64) and the FP64 batch's x of the same,
    widened, arrays."""
    t32 = [torch.from_numpy(np.ascontiguousarray(v, dtype=np.float32)).cuda() for v in (A, B, C, D)]
    X = torch.full((A.size,), -7.0, dtype=torch.float64, device="cuda")
    be.solve_batched_f32(*t32, X, off, n_rows)
    w = [t.double() for t in t32]
    be.solve_batched(*w, off, n_rows)
    torch.cuda.synchronize()
    return X.cpu().numpy(), w[3].cpu().numpy(), be.last_path()


def widened(*arrays):
    return [np.asarray(v, dtype=np.float32).astype(np.float64) for v in arrays]


def system_rows(off, n_rows, size):
    """Mask of the elements that belong to a system (x goes there)."""
    mask = np.zeros(size, dtype=bool)
    for o, n in zip(off, n_rows):
        mask[int(o):int(o) + int(n)] = True
    return mask


def test_batched_f32_matches_oracle_on_widened_values():
    """FP32 storage, FP64 elimination: bit for bit the oracle's solve of the
    exactly widened arrays, and the FP64 batch's; thousands of items, every
    edge length up to 6144 rows."""
    edge = [n for n in EDGE if n <= 6144]
    n_rows = np.concatenate([cases.uneven_lengths(6000, 16, 4096, seed=5), edge,
                             np.arange(1, 16).repeat(50)]).astype(np.int32)
    np.random.default_rng(6).shuffle(n_rows)
    off, n_rows, A, B, C, D = general_uneven(n_rows, seed=61, gap=1)
    A, B, C, D = widened(A, B, C, D)
    note(f"batched f32 {n_rows.size} systems, {int(n_rows.sum())} rows")
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        X, X64, path = run_batched_f32(be, off, n_rows, A, B, C, D)
    finally:
        be.close()
    assert path.startswith("batched"), path
    ref = oracle_solve(off, n_rows, A, B, C, D)
    mask = system_rows(off, n_rows, A.size)
    np.testing.assert_array_equal(X[mask], ref[mask])
    np.testing.assert_array_equal(X[mask], X64[mask])
    assert np.all(X[~mask] == -7.0)  # gaps untouched


@pytest.mark.parametrize("lengths,gap", [
    ([50, 5, 1, 1], 0),   # chunks ending 2, 1 and 0 elements before the arrays' end
    ([97, 6], 0),         # the last quad holds 2 rows: read 2 elements low
    ([300, 7], 0),        # 3 rows: 1 element low
    ([40, 8], 0),         # a full last quad: no shift
    ([1, 2, 1], 1),       # spans of a few elements
    ([99], 0),            # one chunked system, three chunks of 33 rows
])
def test_batched_f32_arrays_end_at_last_row(lengths, gap):
    off, n_rows, A, B, C, D = general_uneven(lengths, seed=len(lengths) + 70, gap=gap)
    end = int(off[-1]) + int(n_rows[-1])
    A, B, C, D = widened(*(v[:end] for v in (A, B, C, D)))
    note(f"batched f32 array end {lengths}")
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        X, X64, _ = run_batched_f32(be, off, n_rows, A, B, C, D)
    finally:
        be.close()
    ref = oracle_solve(off, n_rows, A, B, C, D)
    for o, n in zip(off, n_rows):
        sl = slice(int(o), int(o) + int(n))
        np.testing.assert_array_equal(X[sl], ref[sl])
        np.testing.assert_array_equal(X[sl], X64[sl])


def test_batched_f32_unsupported_layouts():
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        f = torch.zeros(8000, dtype=torch.float32, device="cuda")
        x = torch.zeros(8000, dtype=torch.float64, device="cuda")
        with pytest.raises(FtdError):  # above 128 chunks of 48 rows
            be.solve_batched_f32(f, f, f, f, x, [0], [6145])
        with pytest.raises(FtdError):  # spans fewer than 4 elements
            be.solve_batched_f32(f, f, f, f, x, [0, 1], [1, 2])
    finally:
        be.close()


def test_batched_chunk_32_matches_oracle():
    """32-row chunks (ftd_set_batch_chunk): PaScaL with ceil(n/32) partitions,
    the lane-per-system kernel above 4096 rows."""
    n_rows = np.concatenate([cases.uneven_lengths(3000, 16, 4096, seed=8), EDGE]).astype(np.int32)
    off, n_rows, A, B, C, D = general_uneven(n_rows, seed=81, gap=2)
    note(f"batched chunk 32, {n_rows.size} systems")
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        be.set_batch_chunk(32)
        _, _, X, path = run_batched(be, off, n_rows, A, B, C, D)
        with pytest.raises(FtdError):
            be.set_batch_chunk(40)
    finally:
        be.close()
    assert "chunk=32" in path, path
    np.testing.assert_array_equal(X, oracle_solve(off, n_rows, A, B, C, D, chunk=32))


def test_batched_autotune_keeps_the_faster_chunk():
    n_rows = cases.uneven_lengths(4000, 16, 4096, seed=9)
    off, n_rows, A, B, C, D = general_uneven(n_rows, seed=91, gap=1)
    note(f"batched autotune, {n_rows.size} systems")
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        tA, tB, tC, tD = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (A, B, C, D))
        report = be.autotune_batched(tA, tB, tC, tD, off, n_rows, reps=2)
        torch.cuda.synchronize()
        # the caller's arrays are left as they were
        np.testing.assert_array_equal(tC.cpu().numpy(), C)
        np.testing.assert_array_equal(tD.cpu().numpy(), D)
        best = report["best"]["chunk"]
        assert sorted(c["chunk"] for c in report["candidates"]) == [32, 48]
        assert best in (32, 48)
        be.solve_batched(tA, tB, tC, tD, off, n_rows)
        torch.cuda.synchronize()
        path = be.last_path()
        X = tD.cpu().numpy()
        A32, B32, C32, D32 = widened(A, B, C, D)
        X32, _, path32 = run_batched_f32(be, off, n_rows, A32, B32, C32, D32)
    finally:
        be.close()
    assert f"chunk={best}" in path and f"chunk={best}" in path32, (path, path32)
    np.testing.assert_array_equal(X, oracle_solve(off, n_rows, A, B, C, D, chunk=best))
    mask = system_rows(off, n_rows, A.size)
    ref32 = oracle_solve(off, n_rows, A32, B32, C32, D32, chunk=best)
    np.testing.assert_array_equal(X32[mask], ref32[mask])


def test_batched_toeplitz_c4_shape_sample():
    """The uneven-batch workload's generator (a = c = -r, b = 1 + 2r,
    r ~ U[0.5, 50]) on 2000 systems, back to back."""
    n_rows = cases.uneven_lengths(2000, 16, 4096, seed=0)
    off, n_rows, A, B, C, D, _ = cases.uneven_batch(n_rows, 0.5, 50.0, seed=0)
    note(f"batched toeplitz {n_rows.size} systems, {int(n_rows.sum())} rows")
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        _, _, X, _ = run_batched(be, off, n_rows, A, B, C, D)
    finally:
        be.close()
    np.testing.assert_array_equal(X, oracle_solve(off, n_rows, A, B, C, D))


def test_batched_errors():
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        t = torch.zeros(16, dtype=torch.float64, device="cuda")
        with pytest.raises(FtdError):
            be.solve_batched(t, t, t, t, [0, 4], [4, 0])
        with pytest.raises(FtdError):
            be.solve_batched(t, t, t, t, [-1], [4])
        with pytest.raises(ValueError):  # the layout runs past the arrays
            be.solve_batched(t, t, t, t, [10], [8])
        be.solve_batched(t, t, t, t, [], [])  # empty batch is a no-op
    finally:
        be.close()


@pytest.mark.parametrize("chunk", [32, 48])
def test_batched_line_aligned_split_edges(chunk):
    """The line-aligned split at its edges: systems starting at every offset
    modulo 16 (a gap of 1 between systems walks the residues), with lengths
    just under, at and above the smallest the split takes (chunk + 16 rows
    past a boundary), last chunks of 1..16 rows that borrow 16 from the
    chunk before, and the 128-chunk limit; bit for bit against the oracle
    with the same split, and within rounding of Thomas."""
    base = [chunk + 1, chunk + 15, chunk + 16, chunk + 17, chunk + 31, 2 * chunk + 1,
            2 * chunk + 16, 3 * chunk + 3, 5 * chunk + 15, 5 * chunk + 17, 127 * chunk,
            128 * chunk - 5, 128 * chunk]
    lengths = [n for k in range(16) for n in base[k % len(base):] + base[:k % len(base)]][:160]
    off, n_rows, A, B, C, D = general_uneven(lengths, seed=60 + chunk, gap=1)
    assert len({int(o) % 16 for o in off}) == 16
    note(f"batched line-aligned split edges, chunk {chunk}, {n_rows.size} systems")
    be = TdmaBackendGPU("pascal", 1, 1, 1)
    try:
        be.set_batch_chunk(chunk)
        _, _, X, path = run_batched(be, off, n_rows, A, B, C, D)
    finally:
        be.close()
    assert f"chunk={chunk}" in path, path
    np.testing.assert_array_equal(X, oracle_solve(off, n_rows, A, B, C, D, chunk=chunk))
    T = thomas_all(off, n_rows, A, B, C, D)
    assert np.abs(X - T).max() <= 1e-12 * max(1.0, np.abs(T).max())
