"""Python mirror of the channel pressure solve over the ftd_poisson_* C ABI
(apps/channel_cpu/PressureSolver.cpp:493-533 on one GPU).

    ps = ChannelPoisson(256, 256, 256, Lx=4 * np.pi, Ly=2 * np.pi, Lz=2.0, gamma_z=1.5)
    ps.solve(U, V, W, P, dt)      # device fields with ghosts, [nz+2][ny+2][nx+2]

stream=None runs on torch's current stream (as TdmaBackendGPU does), so
the steps are ordered after the work that wrote the fields there.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .backend import TdmaBackendGPU


def _ptr(x, numel):
    """Device address of x: a contiguous float64 device tensor of `numel`
    elements (checked), or a raw address (the caller's responsibility)."""
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    if hasattr(x, "data_ptr"):
        if not x.is_cuda or not x.is_contiguous() or str(x.dtype) != "torch.float64":
            raise ValueError("expected a contiguous float64 device tensor")
        if x.numel() != numel:
            raise ValueError(f"expected {numel} elements, got {x.numel()}")
        return ctypes.c_void_p(x.data_ptr())
    raise TypeError("expected a device tensor or a device address")


class ChannelPoisson:
    def __init__(self, nx, ny, nz, Lx, Ly, Lz, gamma_z, backend="pascal", partitions=1,
                 eps_constant=0.0, stream=None):
        L = _lib.load()
        kind = int(TdmaBackendGPU.parse(backend)) if isinstance(backend, str) else int(backend)
        h = ctypes.c_void_p()
        st = TdmaBackendGPU._stream(stream)
        _lib.check_poisson(L.ftd_poisson_create(ctypes.byref(h), nx, ny, nz, float(Lx), float(Ly),
                                                float(Lz), float(gamma_z), kind, partitions,
                                                float(eps_constant), st))
        self._h = h
        self.shape = (nz, ny, nx)
        self._cells = nz * ny * nx
        self._field = (nz + 2) * (ny + 2) * (nx + 2)

    def solve(self, U, V, W, P, dt: float) -> None:
        f = self._field
        _lib.check_poisson(_lib.load().ftd_poisson_solve(self._h, _ptr(U, f), _ptr(V, f),
                                                         _ptr(W, f), _ptr(P, f), float(dt)))

    def set_profiling(self, enable: bool) -> None:
        """Time each step (and wait for it); off by default."""
        _lib.check_poisson(_lib.load().ftd_poisson_set_profiling(self._h, int(bool(enable))))

    def solve_rhs(self, rhs, dp) -> None:
        n = self._cells
        _lib.check_poisson(_lib.load().ftd_poisson_solve_rhs(self._h, _ptr(rhs, n), _ptr(dp, n)))

    def set_j_mode(self, mode: int) -> None:
        """J mode of the wall-normal solver (filtered kinds)."""
        _lib.check(_lib.load().ftd_set_j_mode(_lib.load().ftd_poisson_solver(self._h), int(mode)))

    def solver_path(self) -> str:
        return _lib.load().ftd_last_path(_lib.load().ftd_poisson_solver(self._h)).decode()

    def last_ms(self) -> float:
        return _lib.load().ftd_poisson_last_ms(self._h)

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().ftd_poisson_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
