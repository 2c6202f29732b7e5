"""Synthetic tridiagonal batches and the counter-based generator behind them.

Every value is a pure function of (seed, stream, index), so a batch can be
rebuilt bit-for-bit on any host without storing it. The generator is
splitmix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014) applied to a 64-bit counter; the top 53 bits give a
uniform double in [0, 1).

Layout of every array is the reference's: row-major [n_row x n_sys], row j of
system i at j*n_sys + i (libs/pascal_tdma/include/pascal_tdma.hpp:13).
Index order of the RHS counter is system-major: idx = i*n_row + j.
"""
from __future__ import annotations

import numpy as np

_M64 = (1 << 64) - 1
_GAMMA = 0x9E3779B97F4A7C15
_STREAM_MUL = 0xD1B54A32D192ED03


def _splitmix64(z: np.ndarray) -> np.ndarray:
    z = z + np.uint64(_GAMMA)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """U[0,1) doubles for counters `idx` of (seed, stream)."""
    key = np.uint64((seed * _GAMMA + stream * _STREAM_MUL) & _M64)
    with np.errstate(over="ignore"):
        bits = _splitmix64(idx.astype(np.uint64) + key)
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _rhs(seed: int, n_sys: int, n_row: int, lo: float = -1.0, hi: float = 1.0,
         chunk_rows: int = 64) -> np.ndarray:
    """D[j, i] = lo + (hi-lo) * u(seed, 1, i*n_row + j), built in row chunks."""
    D = np.empty((n_row, n_sys), dtype=np.float64)
    cols = np.arange(n_sys, dtype=np.uint64) * np.uint64(n_row)
    for j0 in range(0, n_row, chunk_rows):
        j1 = min(n_row, j0 + chunk_rows)
        idx = cols[None, :] + np.arange(j0, j1, dtype=np.uint64)[:, None]
        D[j0:j1] = lo + (hi - lo) * uniform(seed, 1, idx)
    return D


def toeplitz_batch(n_sys: int, n_row: int, r_lo: float, r_hi: float, seed: int = 0,
                   with_rhs: bool = True):
    """Constant-coefficient batch: a = c = -r_s, b = 1 + 2 r_s per system,
    r_s ~ U[r_lo, r_hi], D ~ U[-1, 1] (rows 0 and n_row-1 included; A[0] and
    C[n_row-1] are present but not read by the non-periodic solvers).

    This is the implicit diffusion operator (1 - r d^2/dz^2) of one ADI sweep;
    rho_s = r_s/(1+2 r_s) < 1/2. Returns (A, B, C, D, r) with A, B, C, D of
    shape [n_row, n_sys] (D is None when with_rhs is False).
    """
    r = r_lo + (r_hi - r_lo) * uniform(seed, 0, np.arange(n_sys, dtype=np.uint64))
    A = np.broadcast_to(-r, (n_row, n_sys)).copy()
    C = A.copy()
    B = np.broadcast_to(1.0 + 2.0 * r, (n_row, n_sys)).copy()
    D = _rhs(seed, n_sys, n_row) if with_rhs else None
    return A, B, C, D, r


def general_batch(n_sys: int, n_row: int, seed: int = 0, margin: float = 0.1):
    """Row-wise diagonally dominant batch with independent random a, b, c:
    a, c ~ U[-1, 1], b = +-(|a| + |c|)(1 + margin + u), D ~ U[-1, 1].
    Exercises every coefficient independently (no Toeplitz structure)."""
    n = n_row * n_sys
    idx = np.arange(n, dtype=np.uint64)
    A = (2.0 * uniform(seed, 2, idx) - 1.0).reshape(n_row, n_sys)
    C = (2.0 * uniform(seed, 3, idx) - 1.0).reshape(n_row, n_sys)
    u = uniform(seed, 4, idx).reshape(n_row, n_sys)
    sgn = np.where(uniform(seed, 5, idx).reshape(n_row, n_sys) < 0.5, -1.0, 1.0)
    B = sgn * (np.abs(A) + np.abs(C)) * (1.0 + margin + u)
    D = _rhs(seed, n_sys, n_row)
    return A, B, C, D


def uneven_lengths(n_systems: int, n_lo: int, n_hi: int, seed: int = 0) -> np.ndarray:
    """N_s = round(exp(U[ln n_lo, ln n_hi])) per system (stream 6)."""
    u = uniform(seed, 6, np.arange(n_systems, dtype=np.uint64))
    n = np.rint(np.exp(np.log(n_lo) + (np.log(n_hi) - np.log(n_lo)) * u)).astype(np.int32)
    return np.clip(n, n_lo, n_hi)


def uneven_batch(n_rows, r_lo: float, r_hi: float, seed: int = 0):
    """Uneven batch in the system-contiguous layout of ftd_solve_batched:
    system s is elements off[s] .. off[s] + n_rows[s] - 1 of the flat arrays
    (systems back to back). a = c = -r_s, b = 1 + 2 r_s, r_s ~ U[r_lo, r_hi]
    (stream 7), D ~ U[-1, 1] over the flat index (stream 8).
    Returns (off, n_rows, A, B, C, D, r)."""
    n_rows = np.asarray(n_rows, dtype=np.int32)
    off = np.zeros(n_rows.size, dtype=np.int64)
    np.cumsum(n_rows[:-1], out=off[1:])
    total = int(n_rows.sum())
    r = r_lo + (r_hi - r_lo) * uniform(seed, 7, np.arange(n_rows.size, dtype=np.uint64))
    rr = np.repeat(r, n_rows)
    A = -rr
    C = -rr
    B = 1.0 + 2.0 * rr
    D = -1.0 + 2.0 * uniform(seed, 8, np.arange(total, dtype=np.uint64))
    return off, n_rows, A, B, C, D, r


def channel_velocity(nx: int, ny: int, nz: int, seed: int):
    """Face velocities U, V, W [nz + 2, ny + 2, nx + 2] ~ N(0, 1) for the
    channel pressure step: periodic ghosts in x and y, no-penetration walls
    (W on the wall faces k = 1 and nz + 1 is zero)."""
    rng = np.random.default_rng(seed)
    sh = (nz + 2, ny + 2, nx + 2)
    U, V, W = (rng.standard_normal(sh) for _ in range(3))
    for F in (U, V, W):
        F[:, :, 0] = F[:, :, nx]
        F[:, :, nx + 1] = F[:, :, 1]
        F[:, 0, :] = F[:, ny, :]
        F[:, ny + 1, :] = F[:, 1, :]
    W[1], W[nz + 1] = 0.0, 0.0
    return U, V, W


def tanh_faces(n: int, gamma: float, L: float) -> np.ndarray:
    """Face positions x_f[0..n] of a tanh-stretched wall-normal grid
    (apps/channel_cpu/Grid.cpp:16-30)."""
    zeta = np.arange(n + 1, dtype=np.float64) / n
    return 0.5 * L * (1.0 - np.tanh(gamma * (1.0 - 2.0 * zeta)) / np.tanh(gamma))


def channel_poisson(n1: int = 128, n2: int = 128, n_row: int = 129, gamma: float = 2.0,
                    Lx: float = 2.0 * np.pi, Ly: float = np.pi, Lz: float = 2.0, seed: int = 0):
    """Wall-normal systems of the channel pressure Poisson solve, one per
    (kx, ky) mode, as PressureSolver::solve_tdma_z_ builds them
    (apps/channel_cpu/PressureSolver.cpp:425-456; wavenumbers :72-78; grid
    spacings Grid.cpp:47-95 with mirrored ghost centres):
      a_k = 1/(dz_k dmz_k) (0 at the lower wall), c_k = 1/(dz_k dmz_{k+1})
      (0 at the upper wall), b_k = -(a_k + c_k) - (kx^2 + ky^2),
    mode s = i + n1 * j, the (0, 0) mode pinned (row 0: a = c = 0, b = 1,
    rhs 0). Right-hand sides are complex, re and im ~ N(0, 1) (Box-Muller on
    streams 9 and 10). Returns (A, B, C, Z) with A, B, C [n_row, n1*n2]
    float64 and Z [n_row, n1*n2] complex128."""
    xf = tanh_faces(n_row, gamma, Lz)
    xc = np.empty(n_row + 2)
    xc[1:n_row + 1] = 0.5 * (xf[:-1] + xf[1:])
    xc[0] = xf[0] - (xc[1] - xf[0])
    xc[n_row + 1] = xf[n_row] + (xf[n_row] - xc[n_row])
    dz = np.empty(n_row + 2)
    dz[1:n_row + 1] = xf[1:] - xf[:-1]
    dz[0], dz[n_row + 1] = dz[1], dz[n_row]
    dmz = np.empty(n_row + 2)
    dmz[1:] = xc[1:] - xc[:-1]
    dmz[0] = dmz[1]
    k = np.arange(n_row)
    a = 1.0 / (dz[k + 1] * dmz[k + 1])
    c = 1.0 / (dz[k + 1] * dmz[k + 2])
    a[0] = 0.0
    c[n_row - 1] = 0.0
    dx, dy = Lx / n1, Ly / n2
    kx2 = 2.0 * (1.0 - np.cos(2.0 * np.pi * np.arange(n1) / n1)) / (dx * dx)
    jj = np.arange(n2)
    jj = np.where(jj <= n2 // 2, jj, jj - n2)
    ky2 = 2.0 * (1.0 - np.cos(2.0 * np.pi * jj / n2)) / (dy * dy)
    lam = (kx2[None, :] + ky2[:, None]).reshape(-1)  # s = i + n1 * j
    n_sys = n1 * n2
    A = np.broadcast_to(a[:, None], (n_row, n_sys)).copy()
    C = np.broadcast_to(c[:, None], (n_row, n_sys)).copy()
    B = -(a + c)[:, None] - lam[None, :]
    idx = np.arange(n_row * n_sys, dtype=np.uint64)
    u1 = np.maximum(uniform(seed, 9, idx), 1e-300)
    u2 = uniform(seed, 10, idx)
    rad = np.sqrt(-2.0 * np.log(u1))
    Z = (rad * np.cos(2.0 * np.pi * u2) + 1j * rad * np.sin(2.0 * np.pi * u2)).reshape(n_row, n_sys)
    A[0, 0] = 0.0
    B[0, 0] = 1.0
    C[0, 0] = 0.0
    Z[0, 0] = 0.0
    return A, np.ascontiguousarray(B), C, np.ascontiguousarray(Z)
