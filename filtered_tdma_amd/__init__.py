"""MI355X (gfx950) batched tridiagonal engine: a drop-in for the
solve(A,B,C,D,n_sys,n_row) path behind dontempty/Filtered_TDMA's
TdmaBackendGPU runtime switch ("pascal" | "filtered" | "filtered_v2").

The product is libftdma.so (HIP kernels + C ABI, include/ftdma.h); this
package holds its sources (csrc/), its build recipe (build.py), the Python
mirror of the reference class (backend.py) and the synthetic workloads
(cases.py).
"""
from ._lib import FtdError
from .backend import Kind, TdmaBackendGPU

__all__ = ["FtdError", "Kind", "TdmaBackendGPU"]
