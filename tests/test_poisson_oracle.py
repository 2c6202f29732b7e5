"""Pins the oracle's channel pressure solve (oracle/poisson_oracle.py)
before it checks ftd_poisson_*: by what the projection must achieve and by
the grid's closed forms, not by retyping its formulas.

  * after one pressure step on a no-penetration velocity field, the
    discrete divergence of the new field vanishes (to rounding), with the
    exact Thomas and with PaScaL's partitions;
  * the pinned (0, 0) mode: the plane mean of dp at the lower wall row is 0;
  * the tanh grid's face spacings sum to Lz, mirror about the midplane, and
    the mirrored ghost centre makes dmz[1] = dz[1];
  * a single Fourier mode in x and y with a smooth wall-normal profile is
    solved with the same wall-normal operator as the matrix form.
"""
import numpy as np
import pytest

import poisson_oracle as PO


@pytest.mark.parametrize("nx,ny,nz,gamma,P", [(16, 12, 17, 2.0, 1), (32, 8, 33, 1.5, 4),
                                              (8, 6, 9, 0.0, 1)])
def test_projection_leaves_no_divergence(nx, ny, nz, gamma, P):
    ch = PO.Channel(nx, ny, nz, 4 * np.pi, 2 * np.pi, 2.0, gamma)
    U, V, W = PO.random_velocity(nx, ny, nz, seed=nx + nz)
    Pf = np.zeros_like(U)
    dt = 1e-3
    div0 = ch.divergence(U, V, W, dt)
    ch.solve(U, V, W, Pf, dt, P=P)
    PO.refresh_ghosts(U, V, nx, ny)
    div1 = ch.divergence(U, V, W, dt)
    assert np.max(np.abs(div1)) <= 1e-10 * np.max(np.abs(div0)), np.max(np.abs(div1))
    # the wall faces are untouched
    assert np.all(W[1] == 0.0) and np.all(W[nz + 1] == 0.0)


def test_zero_mode_pin():
    ch = PO.Channel(16, 8, 17, 2 * np.pi, np.pi, 2.0, 2.0)
    rhs = np.random.default_rng(1).standard_normal((17, 8, 16))
    dp = ch.solve_rhs(rhs)
    assert abs(dp[0].mean()) <= 1e-12 * np.max(np.abs(dp))


def test_tanh_grid_closed_forms():
    for n, g, L in ((17, 2.0, 2.0), (256, 1.5, 2.0), (129, 2.0, 1.0)):
        dz, dmz = PO.axis_metrics(n, 0, g, L, False)
        assert abs(dz[1:n + 1].sum() - L) <= 1e-13 * L
        np.testing.assert_allclose(dz[1:n + 1], dz[n:0:-1], rtol=1e-10)
        assert abs(dmz[1] - dz[1]) <= 1e-15 * L and abs(dmz[n + 1] - dz[n]) <= 1e-15 * L
        # centre spacings sum to the distance between the two ghost centres
        assert abs(dmz[1:n + 2].sum() - (L + 0.5 * (dz[1] + dz[n]))) <= 1e-13 * L
        # the first face spacing follows from the tanh law at zeta = 1/n
        z1 = 0.5 * L * (1 - np.tanh(g * (1 - 2.0 / n)) / np.tanh(g))
        assert abs(dz[1] - z1) <= 1e-15 * L


def test_single_mode_matches_matrix_form():
    """rhs = cos(2 pi x/Lx) f(z): only modes (+-1, 0) are non-zero, and dp
    is cos(2 pi x/Lx) g(z) with (D_zz - kx2) g = f, D_zz the variable-mesh
    operator with Neumann walls, solved densely."""
    nx, ny, nz, Lx, Ly, Lz, gam = 16, 4, 21, 2 * np.pi, np.pi, 2.0, 1.7
    ch = PO.Channel(nx, ny, nz, Lx, Ly, Lz, gam)
    x = (np.arange(nx) + 0.5) * Lx / nx
    f = np.sin(np.linspace(0.3, 2.9, nz))
    rhs = np.cos(2 * np.pi * x / Lx)[None, None, :] * f[:, None, None] * np.ones((1, ny, 1))
    dp = ch.solve_rhs(rhs)
    kx2 = 2.0 * (1.0 - np.cos(2.0 * np.pi / nx)) / (Lx / nx) ** 2
    M = np.zeros((nz, nz))
    for k in range(nz):
        a = 0.0 if k == 0 else 1.0 / (ch.dz[k + 1] * ch.dmz[k + 1])
        c = 0.0 if k == nz - 1 else 1.0 / (ch.dz[k + 1] * ch.dmz[k + 2])
        M[k, k] = -(a + c) - kx2
        if k > 0:
            M[k, k - 1] = a
        if k < nz - 1:
            M[k, k + 1] = c
    g = np.linalg.solve(M, f)
    want = np.cos(2 * np.pi * x / Lx)[None, None, :] * g[:, None, None] * np.ones((1, ny, 1))
    np.testing.assert_allclose(dp, want, rtol=0, atol=1e-12 * np.max(np.abs(g)))
