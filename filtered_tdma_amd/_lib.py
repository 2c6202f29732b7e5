"""Loader of libftdma.so, the HIP engine behind include/ftdma.h.

There is no fallback: if the library is missing or does not load, every
entry point raises. Build it with `python -m filtered_tdma_amd.build`.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libftdma.so")

# every symbol include/ftdma.h declares
EXPORTS = (
    "ftd_kind_from_string", "ftd_create", "ftd_destroy", "ftd_solve", "ftd_solve_nr",
    "ftd_solve_cyclic", "ftd_set_rho", "ftd_set_eps_constant", "ftd_set_eps", "ftd_set_stream",
    "ftd_set_profiling", "ftd_last_J", "ftd_last_gpu_ms", "ftd_last_comm_ms", "ftd_set_tuning",
    "ftd_last_path", "ftd_last_error_string", "ftd_solve_f32c", "ftd_solve_f32",
    "ftd_solve_f32_refine",
    "ftd_solve_batched", "ftd_solve_batched_f32", "ftd_set_batch_chunk", "ftd_autotune_batched",
    "ftd_solve_complex", "ftd_set_j_mode", "ftd_last_J_systems", "ftd_autotune",
    "ftd_solve_static", "ftd_set_guard",
    "ftd_heat_create", "ftd_heat_destroy", "ftd_heat_init", "ftd_heat_step", "ftd_heat_field",
    "ftd_heat_l2", "ftd_heat_last_step_ms", "ftd_heat_last_error_string", "ftd_heat_set_fused",
    "ftd_heat_set_plan", "ftd_heat_last_plan",
    "ftd_poisson_create", "ftd_poisson_destroy", "ftd_poisson_solver", "ftd_poisson_solve",
    "ftd_poisson_solve_rhs", "ftd_poisson_last_ms", "ftd_poisson_last_error_string",
    "ftd_poisson_set_profiling",
)

_lib = None


class FtdError(RuntimeError):
    """A non-zero ftd_status, with the library's error text."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"ftdma status {status}: {msg}")
        self.status = status


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FtdError(-1, f"{LIB_PATH} is not built; run `python -m filtered_tdma_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)
    c_int, c_double = ctypes.c_int, ctypes.c_double
    L.ftd_kind_from_string.argtypes = [ctypes.c_char_p]
    L.ftd_kind_from_string.restype = c_int
    L.ftd_create.argtypes = [ctypes.POINTER(vp), c_int, c_int, c_int, c_int, c_double, vp]
    L.ftd_destroy.argtypes = [vp]
    L.ftd_solve.argtypes = [vp, dp, dp, dp, dp]
    L.ftd_solve_nr.argtypes = [vp, dp, dp, dp, dp, c_int, c_int]
    L.ftd_solve_cyclic.argtypes = [vp, dp, dp, dp, dp]
    L.ftd_set_rho.argtypes = [vp, dp, dp, dp]
    L.ftd_set_eps_constant.argtypes = [vp, c_double]
    L.ftd_set_eps.argtypes = [vp, c_double]
    L.ftd_set_stream.argtypes = [vp, vp]
    L.ftd_set_profiling.argtypes = [vp, c_int]
    L.ftd_last_J.argtypes = [vp, ip, c_int]
    L.ftd_last_gpu_ms.argtypes = [vp]
    L.ftd_last_gpu_ms.restype = c_double
    L.ftd_last_comm_ms.argtypes = [vp]
    L.ftd_last_comm_ms.restype = c_double
    L.ftd_set_tuning.argtypes = [vp, c_int, c_int, c_int, c_int]
    L.ftd_solve_f32c.argtypes = [vp, dp, dp, dp, dp]
    L.ftd_solve_f32.argtypes = [vp, dp, dp, dp, dp, dp]
    L.ftd_solve_f32_refine.argtypes = [vp, dp, dp, dp, dp, dp, c_int]
    L.ftd_solve_batched.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ip, c_int, dp, dp, dp,
                                    dp]
    L.ftd_solve_batched_f32.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ip, c_int, dp, dp,
                                        dp, dp, dp]
    L.ftd_set_batch_chunk.argtypes = [vp, c_int]
    L.ftd_autotune_batched.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ip, c_int, dp, dp,
                                       dp, dp, c_int, ctypes.c_char_p, ctypes.c_size_t]
    L.ftd_solve_complex.argtypes = [vp, dp, dp, dp, dp]
    L.ftd_set_j_mode.argtypes = [vp, c_int]
    L.ftd_set_guard.argtypes = [vp, c_int]
    L.ftd_autotune.argtypes = [vp, dp, dp, dp, dp, c_int, ctypes.c_char_p, ctypes.c_size_t]
    L.ftd_last_J_systems.argtypes = [vp, ip, ctypes.c_int64]
    L.ftd_solve_static.argtypes = [c_int, c_int, c_int, c_int, c_double, c_int, dp, dp, dp, dp,
                                   vp]
    L.ftd_last_path.argtypes = [vp]
    L.ftd_last_path.restype = ctypes.c_char_p
    L.ftd_last_error_string.restype = ctypes.c_char_p
    ip3 = ctypes.POINTER(ctypes.c_int)
    L.ftd_heat_create.argtypes = [ctypes.POINTER(vp), c_int, c_int, c_int, ip3, c_int, ip3,
                                  c_double, vp]
    L.ftd_heat_destroy.argtypes = [vp]
    L.ftd_heat_init.argtypes = [vp]
    L.ftd_heat_step.argtypes = [vp, c_int, c_int]
    L.ftd_heat_field.argtypes = [vp, vp]
    L.ftd_heat_l2.argtypes = [vp, ctypes.POINTER(c_double)]
    L.ftd_heat_set_fused.argtypes = [vp, c_int]
    L.ftd_heat_set_plan.argtypes = [vp, c_int, c_int, c_int]
    L.ftd_heat_last_step_ms.argtypes = [vp]
    L.ftd_heat_last_step_ms.restype = c_double
    L.ftd_heat_last_error_string.restype = ctypes.c_char_p
    L.ftd_heat_last_plan.argtypes = [vp]
    L.ftd_heat_last_plan.restype = ctypes.c_char_p
    L.ftd_poisson_create.argtypes = [ctypes.POINTER(vp), c_int, c_int, c_int, c_double, c_double,
                                     c_double, c_double, c_int, c_int, c_double, vp]
    L.ftd_poisson_destroy.argtypes = [vp]
    L.ftd_poisson_solver.argtypes = [vp]
    L.ftd_poisson_solver.restype = vp
    L.ftd_poisson_solve.argtypes = [vp, dp, dp, dp, dp, c_double]
    L.ftd_poisson_solve_rhs.argtypes = [vp, dp, dp]
    L.ftd_poisson_set_profiling.argtypes = [vp, c_int]
    L.ftd_poisson_last_ms.argtypes = [vp]
    L.ftd_poisson_last_ms.restype = c_double
    L.ftd_poisson_last_error_string.restype = ctypes.c_char_p
    for name in ("ftd_poisson_create", "ftd_poisson_destroy", "ftd_poisson_solve",
                 "ftd_poisson_solve_rhs", "ftd_poisson_set_profiling"):
        getattr(L, name).restype = c_int
    for name in ("ftd_heat_create", "ftd_heat_destroy", "ftd_heat_init", "ftd_heat_step",
                 "ftd_heat_field", "ftd_heat_l2", "ftd_heat_set_fused", "ftd_heat_set_plan"):
        getattr(L, name).restype = c_int
    for name in ("ftd_create", "ftd_destroy", "ftd_solve", "ftd_solve_nr", "ftd_solve_cyclic",
                 "ftd_set_rho", "ftd_set_eps_constant", "ftd_set_eps", "ftd_set_stream",
                 "ftd_set_profiling", "ftd_last_J", "ftd_set_tuning", "ftd_solve_f32c",
                 "ftd_solve_f32", "ftd_solve_f32_refine", "ftd_solve_batched",
                 "ftd_solve_batched_f32",
                 "ftd_set_batch_chunk", "ftd_autotune_batched",
                 "ftd_solve_complex", "ftd_set_j_mode", "ftd_last_J_systems", "ftd_autotune",
                 "ftd_solve_static", "ftd_set_guard"):
        getattr(L, name).restype = c_int
    _lib = L
    return L


def check(status: int) -> None:
    if status != 0:
        raise FtdError(status, load().ftd_last_error_string().decode())


def check_poisson(status: int) -> None:
    if status != 0:
        raise FtdError(status, load().ftd_poisson_last_error_string().decode())


def check_heat(status: int) -> None:
    if status != 0:
        raise FtdError(status, load().ftd_heat_last_error_string().decode())
