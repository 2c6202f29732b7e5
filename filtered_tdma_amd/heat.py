"""Python mirror of the reference's heat ADI driver (SolveTheta,
apps/heat_gpu/solve_theta.cu:391-650) over the ftd_heat_* C ABI.

    sim = SolveTheta(64, 64, 64, parts=(1, 1, 4), backend="filtered", dt=1e-3)
    sim.run(100, use_graph=True)
    print(sim.l2())          # main.cpp:59-81

dt_policy() restates GlobalParams::load's time-step rules (global.cpp:80-97).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .backend import TdmaBackendGPU


def dt_policy(option: str, rho: float, n: int):
    """(dt, Nt) of option 'order' (Nt None for 'strong') for n interior cells."""
    dx = 2.0 / n
    dt = rho / (1.0 - 2.0 * rho) * (2.0 * dx * dx)
    if option != "order":
        return dt, None
    dt_n = rho / (1.0 - 2.0 * rho) * (2.0 * (2.0 / 512.0) * (2.0 / 512.0))
    q = dt_n * 128.0 / dt
    return dt, int(math.floor(q + 0.5))


class SolveTheta:
    def __init__(self, nx, ny, nz, parts=(1, 1, 1), backend="pascal", dt=1e-3,
                 periodic=(0, 0, 0), stream=None):
        L = _lib.load()
        self.shape = (nz + 2, ny + 2, nx + 2)
        kind = int(TdmaBackendGPU.parse(backend)) if isinstance(backend, str) else int(backend)
        pp = (ctypes.c_int * 3)(*parts)
        pe = (ctypes.c_int * 3)(*[int(bool(x)) for x in periodic])
        h = ctypes.c_void_p()
        st = ctypes.c_void_p(0 if stream is None else
                             (stream if isinstance(stream, int) else stream.cuda_stream))
        _lib.check_heat(L.ftd_heat_create(ctypes.byref(h), nx, ny, nz, pp, kind, pe, float(dt),
                                          st))
        self._h = h

    def run(self, nsteps: int, use_graph: bool = False) -> None:
        _lib.check_heat(_lib.load().ftd_heat_step(self._h, int(nsteps), int(bool(use_graph))))

    def set_fused(self, enable: bool) -> None:
        """Fused sweeps (coefficients generated in the kernel) on or off."""
        _lib.check_heat(_lib.load().ftd_heat_set_fused(self._h, int(bool(enable))))

    def set_plan(self, direction: int, rows_per_lane: int = 0, systems_per_tile: int = 0) -> None:
        """Shape of one direction's fused sweep (0 = automatic)."""
        _lib.check_heat(_lib.load().ftd_heat_set_plan(self._h, int(direction), int(rows_per_lane),
                                                      int(systems_per_tile)))

    def reset(self) -> None:
        _lib.check_heat(_lib.load().ftd_heat_init(self._h))

    def field(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        _lib.check_heat(_lib.load().ftd_heat_field(self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def l2(self) -> float:
        v = ctypes.c_double()
        _lib.check_heat(_lib.load().ftd_heat_l2(self._h, ctypes.byref(v)))
        return v.value

    def last_step_ms(self) -> float:
        return _lib.load().ftd_heat_last_step_ms(self._h)

    def last_plan(self) -> str:
        """How the last enqueued step ran each sweep (fused, filtered_fused or
        materialized, with the fused shapes)."""
        return _lib.load().ftd_heat_last_plan(self._h).decode()

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().ftd_heat_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
