"""ctypes front end of the CPU oracle (oracle/ftdma_oracle.c) and of the
reference CPU build in oracle/_ref (when present).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. Product code never imports it.

All arrays are numpy float64 [n_row, n_sys] (C order), solved in place like
the reference: D <- x, A and C clobbered, B untouched. Every function returns
the solved D and leaves the caller's arrays alone unless `inplace=True`.
"""
from __future__ import annotations

import ctypes
import os
import shutil
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libftdma_oracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
MPIEXEC = "/opt/conda/bin/mpiexec"

SEM_GPU, SEM_CPU = 0, 1
J_GPU_RHS, J_CPU_V1, J_CPU_RHS = 0, 1, 2

_lib = None


def build() -> None:
    """Compile the oracle library with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "libftdma_oracle.so"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        c_int, c_double = ctypes.c_int, ctypes.c_double
        L.ftdo_thomas.argtypes = [dp, dp, dp, dp, c_int, c_int, c_int, c_int]
        L.ftdo_thomas_cyclic.argtypes = [dp, dp, dp, dp, c_int, c_int, c_int, c_int]
        L.ftdo_pascal.argtypes = [dp, dp, dp, dp, c_int, c_int, c_int, c_int, c_int, c_int]
        L.ftdo_pascal_nested.argtypes = [dp, dp, dp, dp, c_int, c_int, c_int, c_int, c_int,
                                         c_int]
        L.ftdo_pascal_starts.argtypes = [dp, dp, dp, dp, c_int, c_int, c_int, ip, c_int, c_int,
                                         c_int]
        L.ftdo_filtered_guard.argtypes = [dp, dp, dp, dp, c_int, c_int, c_int, c_int, ip,
                                          c_double, c_int, ip, c_int]
        L.ftdo_set_rho.argtypes = [dp, dp, dp, c_int, c_int, dp, dp]
        L.ftdo_set_rho.restype = None
        L.ftdo_cal_J.argtypes = [c_double, c_double, c_double, c_int, c_int]
        L.ftdo_filtered.argtypes = [dp, dp, dp, dp, c_int, c_int, c_int, c_int, c_int, c_int,
                                    dp, dp, c_double, c_int, ip, ip, c_int]
        L.ftdo_para_range.argtypes = [c_int, c_int, c_int, c_int, ip, ip]
        L.ftdo_para_range.restype = None
        L.ftdo_last_error.restype = ctypes.c_char_p
        fp = ctypes.POINTER(ctypes.c_float)
        L.ftdo_filtered_learned.argtypes = [dp, dp, dp, dp, c_int, c_int, c_int, c_int,
                                            c_double, c_int, fp, c_int, c_int, ip, c_int]
        L.ftdo_mlp_J.argtypes = [fp, c_int, c_int, fp]
        L.ftdo_pascal_ir.argtypes = [fp, fp, fp, fp, dp, c_int, c_int, c_int, c_int, c_int]
        L.ftdo_lg.argtypes = [c_double]
        L.ftdo_lg.restype = ctypes.c_float
        for f in (L.ftdo_thomas, L.ftdo_thomas_cyclic, L.ftdo_pascal, L.ftdo_pascal_nested,
                  L.ftdo_pascal_starts, L.ftdo_filtered_guard, L.ftdo_pascal_ir,
                  L.ftdo_cal_J,
                  L.ftdo_filtered, L.ftdo_filtered_learned, L.ftdo_mlp_J):
            f.restype = c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _prep(A, B, C, D, inplace):
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (A, B, C, D)]
    if not inplace:
        arrs = [x.copy() for x in arrs]
    n_row, n_sys = arrs[3].shape
    for x in arrs:
        if x.shape != (n_row, n_sys):
            raise ValueError("A, B, C, D must share one [n_row, n_sys] shape")
    return arrs, n_sys, n_row


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().ftdo_last_error().decode())


def thomas(A, B, C, D, sem=SEM_GPU, cyclic=False, nthreads=0, inplace=False):
    (A, B, C, D), n_sys, n_row = _prep(A, B, C, D, inplace)
    f = lib().ftdo_thomas_cyclic if cyclic else lib().ftdo_thomas
    _check(f(_p(A), _p(B), _p(C), _p(D), n_sys, n_row, sem, nthreads))
    return D


def pascal(A, B, C, D, P, sem=SEM_GPU, cyclic=False, nthreads=0, inplace=False):
    (A, B, C, D), n_sys, n_row = _prep(A, B, C, D, inplace)
    _check(lib().ftdo_pascal(_p(A), _p(B), _p(C), _p(D), n_sys, n_row, P, int(cyclic), sem,
                             nthreads))
    return D


def pascal_nested(A, B, C, D, P, G, sem=SEM_GPU, nthreads=0, inplace=False):
    """PaScaL with P partitions whose 2P-row reduced system is solved by
    PaScaL over G groups of rows (G = 1: Thomas, the same as pascal)."""
    (A, B, C, D), n_sys, n_row = _prep(A, B, C, D, inplace)
    _check(lib().ftdo_pascal_nested(_p(A), _p(B), _p(C), _p(D), n_sys, n_row, P, G, sem,
                                    nthreads))
    return D


def pascal_starts(A, B, C, D, starts, G, sem=SEM_GPU, nthreads=0, inplace=False):
    """pascal_nested over the partitions starts[p] .. starts[p + 1] - 1
    (starts from 0 to n_row) instead of the para_range split."""
    (A, B, C, D), n_sys, n_row = _prep(A, B, C, D, inplace)
    st = np.ascontiguousarray(starts, dtype=np.int32)
    _check(lib().ftdo_pascal_starts(_p(A), _p(B), _p(C), _p(D), n_sys, n_row, st.size - 1,
                                    st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), G, sem,
                                    nthreads))
    return D


def pascal_ir(A, B, C, D, Q, sweeps, nthreads=0):
    """FP32 elimination with FP64 refinement: A, B, C, D are cast to float32
    (the stored values); returns x (float64) after the FP32 solve and
    `sweeps` refinement sweeps, PaScaL over Q partitions."""
    A, B, C, D = (np.ascontiguousarray(x, dtype=np.float32) for x in (A, B, C, D))
    n_row, n_sys = A.shape
    X = np.zeros((n_row, n_sys), dtype=np.float64)
    fp = ctypes.POINTER(ctypes.c_float)
    _check(lib().ftdo_pascal_ir(A.ctypes.data_as(fp), B.ctypes.data_as(fp), C.ctypes.data_as(fp),
                                D.ctypes.data_as(fp), _p(X), n_sys, n_row, Q, sweeps, nthreads))
    return X


def filtered_guard(A, B, C, D, P, J, eps, eps_direct=False, cyclic=False, nthreads=0,
                   inplace=False):
    """The guard mode: J [P, n_sys] int (the J of each (partition, system)
    before the guard); returns (x, records [P, n_sys] = J used | guard << 17).
    cyclic: periodic systems (wrapped interfaces, periodic reduced system)."""
    (A, B, C, D), n_sys, n_row = _prep(A, B, C, D, inplace)
    Jm = np.ascontiguousarray(np.broadcast_to(np.asarray(J, dtype=np.int32), (P, n_sys)))
    rec = np.zeros((P, n_sys), dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    _check(lib().ftdo_filtered_guard(_p(A), _p(B), _p(C), _p(D), n_sys, n_row, P, int(bool(cyclic)),
                                     Jm.ctypes.data_as(ip), float(eps), int(bool(eps_direct)),
                                     rec.ctypes.data_as(ip), nthreads))
    return D, rec


def set_rho(A, B, C):
    n_row, n_sys = np.shape(A)
    A, B, C = (np.ascontiguousarray(x, dtype=np.float64) for x in (A, B, C))
    ar = np.zeros(n_row + 1)
    cr = np.zeros(n_row + 1)
    lib().ftdo_set_rho(_p(A), _p(B), _p(C), n_sys, n_row, _p(ar), _p(cr))
    return ar, cr


def cal_J(rho, bound, eps, n_row, which=J_GPU_RHS):
    return lib().ftdo_cal_J(float(rho), float(bound), float(eps), int(n_row), int(which))


def filtered(A, B, C, D, P, version, eps, eps_direct=False, sem=SEM_GPU, cyclic=False,
             rho=None, J_force=None, nthreads=0, inplace=False, return_J=False):
    """rho: None (reference state before any set_rho: all zero -> full J),
    'auto' (column-0 bounds of these A, B, C, as set_rho_device), or a
    (A_rho, C_rho) pair of length n_row+1."""
    if isinstance(rho, str) and rho == "auto":
        rho = set_rho(A, B, C)
    (A, B, C, D), n_sys, n_row = _prep(A, B, C, D, inplace)
    ar = cr = None
    if rho is not None:
        ar = np.ascontiguousarray(rho[0], dtype=np.float64)
        cr = np.ascontiguousarray(rho[1], dtype=np.float64)
    Jout = np.zeros(P, dtype=np.int32)
    jf = None
    if J_force is not None:
        jf = np.ascontiguousarray(np.broadcast_to(np.asarray(J_force, dtype=np.int32), (P,)))
    ip = ctypes.POINTER(ctypes.c_int)
    _check(lib().ftdo_filtered(
        _p(A), _p(B), _p(C), _p(D), n_sys, n_row, P, version, int(cyclic), sem,
        _p(ar), _p(cr), float(eps), int(bool(eps_direct)),
        jf.ctypes.data_as(ip) if jf is not None else None, Jout.ctypes.data_as(ip), nthreads))
    return (D, Jout) if return_J else D


MLP_JSON = os.path.join(os.path.dirname(HERE), "filtered_tdma_amd", "learned_j", "mlp_j.json")


def mlp_weights(path=MLP_JSON):
    """(flat float32 weights W1, b1, W2, b2, W3, b3; hidden width H; margin)
    of the J network."""
    import json

    with open(path) as f:
        w = json.load(f)
    flat = np.concatenate([np.asarray(w["W1"], np.float32).reshape(-1),
                           np.asarray(w["b1"], np.float32), np.asarray(w["W2"], np.float32).reshape(-1),
                           np.asarray(w["b2"], np.float32), np.asarray(w["W3"], np.float32),
                           np.asarray([w["b3"]], np.float32)])
    H = int(w["hidden"])
    assert flat.size == H * H + 8 * H + 1
    return np.ascontiguousarray(flat), H, int(w["margin"])


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def mlp_J(features, weights=None):
    """The network's J (before the clamp to [0, m-1]) for 5 features."""
    w, H, margin = weights if weights is not None else mlp_weights()
    f = np.ascontiguousarray(features, dtype=np.float32)
    return int(lib().ftdo_mlp_J(_fp(w), H, margin, _fp(f)))


def lg(x):
    return float(lib().ftdo_lg(float(x)))


def filtered_learned(A, B, C, D, P, eps, eps_direct=False, cyclic=False, weights=None,
                     nthreads=0, inplace=False):
    """Learned-J filtered solve (v1 sweeps, GPU semantics): returns x and the
    J per (partition, system) with bit 16 = eps_ fallback, shape [P, n_sys]."""
    w, H, margin = weights if weights is not None else mlp_weights()
    (A, B, C, D), n_sys, n_row = _prep(A, B, C, D, inplace)
    Jout = np.zeros(P * n_sys, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    _check(lib().ftdo_filtered_learned(
        _p(A), _p(B), _p(C), _p(D), n_sys, n_row, P, int(cyclic), float(eps),
        int(bool(eps_direct)), _fp(w), int(H), int(margin), Jout.ctypes.data_as(ip), nthreads))
    return D, Jout.reshape(P, n_sys)


def para_range(start, end, nproc, rank):
    a, b = ctypes.c_int(), ctypes.c_int()
    lib().ftdo_para_range(start, end, nproc, rank, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


# --------------------------- reference CPU build ---------------------------

def ref_available() -> bool:
    return os.path.exists(REF_DRIVER) and os.path.exists(MPIEXEC)


_KIND = {"filtered": 0, "filtered_v2": 1, "pascal": 2}


def ref_run(mode, kind, A, B, C, D, P, cyclic=False, do_rho=True, eps_constant=0.0,
            J=None, timeout=300):
    """Solve with the reference's own CPU code (oracle/_ref/ref_driver).
    mode 'backend' runs TdmaBackend under mpiexec -n P; mode 'steps' runs the
    filtered step functions with forced J in one process."""
    if not ref_available():
        raise RuntimeError("reference CPU build missing: make -C oracle -f Makefile.ref")
    n_row, n_sys = np.shape(D)
    Jv = np.zeros(P, dtype=np.int32) if J is None else np.broadcast_to(
        np.asarray(J, dtype=np.int32), (P,)).copy()
    tmp = tempfile.mkdtemp(prefix="ftdref_")
    try:
        fin = os.path.join(tmp, "in.bin")
        fout = os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            f.write(b"FTDI")
            f.write(struct.pack("<6i", n_sys, n_row, P, _KIND[kind], int(cyclic), int(do_rho)))
            f.write(struct.pack("<d", float(eps_constant)))
            f.write(Jv.astype("<i4").tobytes())
            for x in (A, B, C, D):
                f.write(np.ascontiguousarray(x, dtype="<f8").tobytes())
        if mode == "backend":
            cmd = [MPIEXEC, "-n", str(P), REF_DRIVER, "backend", fin, fout]
        else:
            cmd = [REF_DRIVER, "steps", fin, fout]
        subprocess.run(cmd, check=True, timeout=timeout, capture_output=True)
        return np.fromfile(fout, dtype="<f8").reshape(n_row, n_sys)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


# ------------------------------------------------------------- heat harness

def heat_dt(option, rho, n):
    """dt policy of GlobalParams::load (global.cpp:80-97) for n interior
    cells: (dt, Nt) for option 'order'; dt only for 'strong'."""
    dx = 2.0 / n
    dt = rho / (1.0 - 2.0 * rho) * (2.0 * dx * dx)
    if option == "order":
        dt_n = rho / (1.0 - 2.0 * rho) * (2.0 * (2.0 / 512.0) * (2.0 / 512.0))
        tmax = dt_n * 128.0
        import math
        q = tmax / dt
        nt = int(math.floor(q + 0.5)) if q >= 0 else -int(math.floor(-q + 0.5))
        return dt, nt
    return dt, None


def heat_run(nx, ny, nz, parts, kind, dt, Nt, periodic=(0, 0, 0), sem=SEM_CPU, nthreads=0,
             return_theta=False):
    """Run the heat ADI restatement on a parts = (px, py, pz) grid (an int
    means (1, 1, pz)); returns L2 (and the field with ghosts)."""
    L = lib()
    if isinstance(parts, int):
        parts = (1, 1, parts)
    if not hasattr(L, "_heat_ready"):
        L.ftdo_heat_run.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int),
                                    ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                    ctypes.c_double, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.ftdo_heat_run.restype = ctypes.c_int
        L._heat_ready = True
    kinds = {"filtered": 0, "filtered_v2": 1, "pascal": 2}
    per = (ctypes.c_int * 3)(*periodic)
    th = np.zeros((nz + 2, ny + 2, nx + 2)) if return_theta else None
    l2 = ctypes.c_double()
    pp = (ctypes.c_int * 3)(*parts)
    rc = L.ftdo_heat_run(nx, ny, nz, pp, kinds[kind], sem, per, float(dt), int(Nt),
                         _p(th) if th is not None else None, ctypes.byref(l2), nthreads)
    if rc != 0:
        raise RuntimeError(f"heat_run failed ({rc}): " + L.ftdo_last_error().decode())
    return (l2.value, th) if return_theta else l2.value
