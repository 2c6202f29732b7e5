"""Python mirror of the reference's TdmaBackendGPU
(apps/heat_gpu/tdma_backend_gpu.hpp:21-70) over the C ABI of libftdma.so.

Same names and meaning as the reference class: Kind, parse, solve,
solve_cyclic, set_rho_device, set_eps_constant, last_comm_ms, last_gpu_ms,
plus last_J (FilteredTDMACUDA::last_J, filtered_tdma_cuda.cu:83-87). Errors
raise FtdError instead of calling exit(1).

Arrays are anything exposing a device pointer: torch CUDA tensors (float64,
contiguous, shape [n_row, n_sys]) or raw integer addresses. torch is used as
plumbing for device memory and streams only; every solve runs the HIP
kernels of libftdma.so.
"""
from __future__ import annotations

import ctypes
import enum

from . import _lib


class Kind(enum.IntEnum):
    """TdmaBackendGPU::Kind (tdma_backend_gpu.hpp:23)."""
    FILTERED = 0
    FILTERED_V2 = 1
    PASCAL = 2


def _ptr(x, n_row=None, n_sys=None, dtype="float64"):
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    import torch

    if not isinstance(x, torch.Tensor):
        raise TypeError("arrays must be torch tensors or integer device addresses")
    want = getattr(torch, dtype)
    if not x.is_cuda or x.dtype != want or not x.is_contiguous():
        raise ValueError(f"arrays must be contiguous {dtype} device tensors")
    if n_row is not None and tuple(x.shape) != (n_row, n_sys):
        raise ValueError(f"expected shape ({n_row}, {n_sys}), got {tuple(x.shape)}")
    return ctypes.c_void_p(x.data_ptr())


class TdmaBackendGPU:
    Kind = Kind

    @staticmethod
    def parse(s: str) -> Kind:
        """tdma_backend_gpu.cu:6-13: case-insensitive, unknown -> FILTERED."""
        return Kind(_lib.load().ftd_kind_from_string(None if s is None else s.encode()))

    def __init__(self, kind, n_sys: int, n_row: int, partitions: int = 1,
                 eps_constant: float = 0.0, stream=None):
        L = _lib.load()
        if isinstance(kind, str):
            kind = self.parse(kind)
        self.kind = Kind(kind)
        self.n_sys, self.n_row, self.partitions = int(n_sys), int(n_row), int(partitions)
        h = ctypes.c_void_p()
        _lib.check(L.ftd_create(ctypes.byref(h), int(self.kind), self.n_sys, self.n_row,
                                self.partitions, float(eps_constant), self._stream(stream)))
        self._h = h

    @staticmethod
    def _stream(stream):
        if stream is None:
            try:
                import torch

                if torch.cuda.is_available():
                    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
            except ImportError:
                pass
            return ctypes.c_void_p(0)
        if isinstance(stream, int):
            return ctypes.c_void_p(stream)
        return ctypes.c_void_p(stream.cuda_stream)

    def _args(self, A, B, C, D, n_row=None, n_sys=None):
        n_row = self.n_row if n_row is None else n_row
        n_sys = self.n_sys if n_sys is None else n_sys
        return [_ptr(x, n_row, n_sys) for x in (A, B, C, D)]

    def solve(self, A, B, C, D) -> None:
        _lib.check(_lib.load().ftd_solve(self._h, *self._args(A, B, C, D)))

    def solve_nr(self, A, B, C, D, n_sys: int, n_row: int) -> None:
        _lib.check(_lib.load().ftd_solve_nr(self._h, *self._args(A, B, C, D, n_row, n_sys),
                                            int(n_sys), int(n_row)))

    def solve_cyclic(self, A, B, C, D) -> None:
        _lib.check(_lib.load().ftd_solve_cyclic(self._h, *self._args(A, B, C, D)))

    def solve_f32c(self, A, B, C, D) -> None:
        """FP32 A, B, C; FP64 D <- x (28 B/row)."""
        a, b, c = (_ptr(x, self.n_row, self.n_sys, "float32") for x in (A, B, C))
        _lib.check(_lib.load().ftd_solve_f32c(self._h, a, b, c, _ptr(D, self.n_row, self.n_sys)))

    def solve_f32(self, A, B, C, D, X) -> None:
        """FP32 A, B, C, D; x to FP64 X (24 B/row)."""
        a, b, c, d = (_ptr(x, self.n_row, self.n_sys, "float32") for x in (A, B, C, D))
        _lib.check(_lib.load().ftd_solve_f32(self._h, a, b, c, d, _ptr(X, self.n_row, self.n_sys)))

    def solve_f32_refine(self, A, B, C, D, X, sweeps: int) -> None:
        """FP32 A, B, C, D; FP32 elimination, FP64 reduced solve and update,
        then `sweeps` FP64-residual refinement sweeps; x to FP64 X
        (24 + 32 sweeps B/row)."""
        a, b, c, d = (_ptr(x, self.n_row, self.n_sys, "float32") for x in (A, B, C, D))
        _lib.check(_lib.load().ftd_solve_f32_refine(self._h, a, b, c, d,
                                                    _ptr(X, self.n_row, self.n_sys),
                                                    int(sweeps)))

    def solve_complex(self, A, B, C, Z) -> None:
        """Complex right-hand sides: Z complex128 [n_row, n_sys] (or float64
        [n_row, 2 n_sys], (re, im) interleaved) <- x; A, B, C real."""
        import torch

        a, b, c = self._args(A, B, C, A)[:3]
        if isinstance(Z, int):
            z = ctypes.c_void_p(Z)
        else:
            shape = (self.n_row, self.n_sys) if Z.dtype == torch.complex128 else \
                (self.n_row, 2 * self.n_sys)
            if Z.dtype not in (torch.complex128, torch.float64) or not Z.is_cuda or \
                    not Z.is_contiguous() or tuple(Z.shape) != shape:
                raise ValueError(f"Z must be a contiguous complex128 {shape[0]}x{self.n_sys} "
                                 "device tensor (or its float64 view)")
            z = ctypes.c_void_p(Z.data_ptr())
        _lib.check(_lib.load().ftd_solve_complex(self._h, a, b, c, z))

    def set_j_mode(self, mode: int) -> None:
        """0: the reference's cal_J per partition; 1: learned J per system
        (FILTERED kind), checked against eps_."""
        _lib.check(_lib.load().ftd_set_j_mode(self._h, int(mode)))

    def set_guard(self, enable: bool) -> None:
        """Filtered kinds: solve exactly the systems where the terms the
        filtered solve drops exceed eps_/4 (v1 sweeps for every kind)."""
        _lib.check(_lib.load().ftd_set_guard(self._h, int(bool(enable))))

    def last_J_records(self, complex_rhs: bool = False):
        """The per (partition, system) records of the last learned or guarded
        solve, shape [P, n_sys] (n_sys doubled for a complex solve):
        J | fallback << 16 | guard << 17."""
        import numpy as np

        n = self.partitions * self.n_sys * (2 if complex_rhs else 1)
        out = np.empty(n, dtype=np.int32)
        _lib.check(_lib.load().ftd_last_J_systems(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n))
        return out.reshape(self.partitions, -1)

    def last_J_systems(self, complex_rhs: bool = False):
        """(J, fallback) arrays of shape [P, n_sys] (n_sys doubled for a
        complex solve) from the last learned-mode solve."""
        out = self.last_J_records(complex_rhs)
        return out & 0xFFFF, (out >> 16) & 1

    def last_J(self, n=None) -> list:
        n = self.partitions if n is None else int(n)
        out = (ctypes.c_int * n)()
        _lib.check(_lib.load().ftd_last_J(self._h, out, n))
        return list(out)

    @staticmethod
    def _layout(row_offsets, n_rows, arrays, names):
        """Host layout arrays, checked against the device arrays' lengths."""
        import numpy as np

        off = np.ascontiguousarray(row_offsets, dtype=np.int64)
        rows = np.ascontiguousarray(n_rows, dtype=np.int32)
        if off.shape != rows.shape or off.ndim != 1:
            raise ValueError("row_offsets and n_rows must be 1-D of one length")
        if off.size:
            need = int(np.max(off + rows))
            for name, x in zip(names, arrays):
                if hasattr(x, "numel") and x.numel() < need:
                    raise ValueError(f"{name} holds {x.numel()} elements; the layout reaches {need}")
        return (off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(off.size), off, rows)

    def solve_batched(self, A, B, C, D, row_offsets, n_rows) -> None:
        """Uneven batch (no reference counterpart): system s is elements
        row_offsets[s] .. row_offsets[s] + n_rows[s] - 1 of the 1-D float64
        device arrays A, B, C, D; D <- x. row_offsets (int64) and n_rows
        (int32) are host arrays."""
        po, pr, n, _, _ = self._layout(row_offsets, n_rows, (A, B, C, D), "ABCD")
        ptrs = [_ptr(x) for x in (A, B, C, D)]
        _lib.check(_lib.load().ftd_solve_batched(self._h, po, pr, n, *ptrs))

    def set_batch_chunk(self, rows: int) -> None:
        """Rows per chunk of the batched plans: 32, 48, or 0 for the default."""
        _lib.check(_lib.load().ftd_set_batch_chunk(self._h, int(rows)))

    def autotune_batched(self, A, B, C, D, row_offsets, n_rows, reps: int = 5) -> dict:
        """Times the batched solve of this layout with each chunk length (on
        copies of C and D) and keeps the faster; returns the report."""
        import json

        po, pr, n, _, _ = self._layout(row_offsets, n_rows, (A, B, C, D), "ABCD")
        ptrs = [_ptr(x) for x in (A, B, C, D)]
        buf = ctypes.create_string_buffer(2048)
        _lib.check(_lib.load().ftd_autotune_batched(self._h, po, pr, n, *ptrs, int(reps), buf,
                                                    len(buf)))
        return json.loads(buf.value.decode())

    def solve_batched_f32(self, A, B, C, D, X, row_offsets, n_rows) -> None:
        """The uneven batch with FP32 storage: A, B, C, D float32, the
        elimination in FP64 on their exactly widened values, x into the
        float64 array X at the same offsets (systems of up to 6144 rows)."""
        po, pr, n, _, _ = self._layout(row_offsets, n_rows, (A, B, C, D, X), "ABCDX")
        ptrs = [_ptr(x, dtype="float32") for x in (A, B, C, D)] + [_ptr(X)]
        _lib.check(_lib.load().ftd_solve_batched_f32(self._h, po, pr, n, *ptrs))

    def set_rho_device(self, A, B, C) -> None:
        a, b, c, _ = self._args(A, B, C, A)
        _lib.check(_lib.load().ftd_set_rho(self._h, a, b, c))

    def set_eps_constant(self, eps_constant: float) -> None:
        _lib.check(_lib.load().ftd_set_eps_constant(self._h, float(eps_constant)))

    def set_eps(self, eps: float) -> None:
        _lib.check(_lib.load().ftd_set_eps(self._h, float(eps)))

    def set_stream(self, stream) -> None:
        _lib.check(_lib.load().ftd_set_stream(self._h, self._stream(stream)))

    def set_profiling(self, enable: bool) -> None:
        _lib.check(_lib.load().ftd_set_profiling(self._h, int(bool(enable))))

    def set_tuning(self, rows_per_lane=0, max_threads=0, systems_per_tile=0,
                   lanes_per_system=0) -> None:
        _lib.check(_lib.load().ftd_set_tuning(self._h, rows_per_lane, max_threads,
                                              systems_per_tile, lanes_per_system))

    def autotune(self, A, B, C, D, reps: int = 3) -> dict:
        """Time every candidate kernel shape on copies of D and keep the
        fastest for later solves; returns the candidates' times and the
        choice (the library's JSON report)."""
        import json

        buf = ctypes.create_string_buffer(1 << 16)
        _lib.check(_lib.load().ftd_autotune(self._h, *self._args(A, B, C, D), int(reps), buf,
                                            len(buf)))
        return json.loads(buf.value.decode())

    def last_gpu_ms(self) -> float:
        return _lib.load().ftd_last_gpu_ms(self._h)

    def last_comm_ms(self) -> float:
        return _lib.load().ftd_last_comm_ms(self._h)

    def last_path(self) -> str:
        return _lib.load().ftd_last_path(self._h).decode()

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().ftd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
