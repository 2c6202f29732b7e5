"""The channel pressure solve on the GPU (ftd_poisson_*) against the
oracle's restatement (oracle/poisson_oracle.py). The FFTs are hipFFT's
here and pocketfft's there, so the comparison is within 1e-11 of the
fields' scale (the wall-normal solves themselves are bit-identical
partitions); the projected field's divergence must vanish either way."""
import sys

import numpy as np
import pytest

import poisson_oracle as PO
from filtered_tdma_amd.poisson import ChannelPoisson

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


CASES = [
    (32, 16, 33, 2.0, 1),     # wall-normal Thomas in one lane
    (64, 32, 65, 1.5, 4),     # PaScaL, 4 partitions
    (16, 8, 17, 0.0, 1),      # uniform wall-normal grid
    (48, 20, 129, 2.0, 1),    # the channel's n_row, internal partitions
]


@pytest.mark.parametrize("nx,ny,nz,gamma,P", CASES)
def test_pressure_step_matches_oracle(nx, ny, nz, gamma, P):
    note(f"poisson step nx={nx} ny={ny} nz={nz} gamma={gamma} P={P}")
    Lx, Ly, Lz, dt = 4 * np.pi, 2 * np.pi, 2.0, 2e-3
    U, V, W = PO.random_velocity(nx, ny, nz, seed=nx * ny + nz)
    Pf = np.random.default_rng(3).standard_normal(U.shape)
    dev = [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in (U, V, W, Pf)]
    ps = ChannelPoisson(nx, ny, nz, Lx, Ly, Lz, gamma, "pascal", P)
    try:
        ps.set_profiling(True)
        ps.solve(*dev, dt)
        torch.cuda.synchronize()
        assert ps.last_ms() > 0.0
    finally:
        ps.close()
    got = [d.cpu().numpy() for d in dev]
    ch = PO.Channel(nx, ny, nz, Lx, Ly, Lz, gamma)
    ref = [f.copy() for f in (U, V, W, Pf)]
    ch.solve(*ref, dt, P=P)
    for name, g, r, f0 in zip("UVWP", got, ref, (U, V, W, Pf)):
        scale = max(np.max(np.abs(r - f0)), 1e-300)
        err = np.max(np.abs(g - r))
        assert err <= 1e-11 * scale + 1e-14 * np.max(np.abs(r)), (name, err, scale)
        # ghosts are not the step's to write
        np.testing.assert_array_equal(g[0], f0[0])
    Ug, Vg, Wg = got[0].copy(), got[1].copy(), got[2]
    PO.refresh_ghosts(Ug, Vg, nx, ny)
    div0 = ch.divergence(U, V, W, dt)
    div1 = ch.divergence(Ug, Vg, Wg, dt)
    assert np.max(np.abs(div1)) <= 1e-10 * np.max(np.abs(div0))


def test_poisson_rhs_only():
    nx, ny, nz = 32, 24, 41
    note(f"poisson rhs nx={nx} ny={ny} nz={nz}")
    rhs = np.random.default_rng(7).standard_normal((nz, ny, nx))
    ps = ChannelPoisson(nx, ny, nz, 2 * np.pi, np.pi, 2.0, 2.0)
    try:
        t_rhs = torch.from_numpy(rhs).cuda()
        t_dp = torch.empty_like(t_rhs)
        ps.solve_rhs(t_rhs, t_dp)
        dp = t_dp.cpu().numpy()
        np.testing.assert_array_equal(t_rhs.cpu().numpy(), rhs)  # input kept
        with pytest.raises(ValueError):  # a field of the wrong size never reaches the GPU
            ps.solve_rhs(t_rhs[:-1], t_dp)
    finally:
        ps.close()
    ref = PO.Channel(nx, ny, nz, 2 * np.pi, np.pi, 2.0, 2.0).solve_rhs(rhs)
    assert np.max(np.abs(dp - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_steps_ordered_after_torch_work():
    """Fields restored on torch's stream (asynchronous device copies) right
    before each step: the handle runs on that stream by default, so every
    step sees the restored fields."""
    nx, ny, nz, dt = 256, 256, 256, 1e-3
    note(f"poisson ordering nx={nx} ny={ny} nz={nz}")
    U, V, W = PO.random_velocity(nx, ny, nz, seed=5)
    Pf = np.zeros_like(U)
    dev0 = [torch.from_numpy(f).cuda() for f in (U, V, W, Pf)]
    dev = [torch.empty_like(d) for d in dev0]
    ps = ChannelPoisson(nx, ny, nz, 4 * np.pi, 2 * np.pi, 2.0, 1.5)
    ch = PO.Channel(nx, ny, nz, 4 * np.pi, 2 * np.pi, 2.0, 1.5)
    div0 = np.max(np.abs(ch.divergence(U, V, W, dt)))
    try:
        for _ in range(3):
            for d, d0 in zip(dev, dev0):
                d.copy_(d0)
            ps.solve(*dev, dt)
            got = [d.cpu().numpy() for d in dev[:3]]
            PO.refresh_ghosts(got[0], got[1], nx, ny)
            assert np.max(np.abs(ch.divergence(*got, dt))) <= 1e-10 * div0
    finally:
        ps.close()
