"""ftd_solve_static, the entry point that allocates nothing, against the
handle entry points (ftd_create + ftd_set_rho + ftd_solve / ftd_solve_cyclic)
through the C ABI: bit-identical x for each kind and path, on two calls in a
row per case, the second with a smaller input (the static scratch array
keeps the first call's contents)."""
import ctypes
import sys

import numpy as np
import pytest

from filtered_tdma_amd import _lib, cases

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KINDS = {"filtered": 0, "filtered_v2": 1, "pascal": 2}


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def with_handle(L, kind, A, B, C, D, P, eps_c, cyclic, stream):
    n_row, n_sys = D.shape
    tA, tB, tC, tD = dev(A), dev(B), dev(C), dev(D)
    h = ctypes.c_void_p()
    _lib.check(L.ftd_create(ctypes.byref(h), KINDS[kind], n_sys, n_row, P, eps_c, stream))
    try:
        _lib.check(L.ftd_set_rho(h, ptr(tA), ptr(tB), ptr(tC)))
        f = L.ftd_solve_cyclic if cyclic else L.ftd_solve
        _lib.check(f(h, ptr(tA), ptr(tB), ptr(tC), ptr(tD)))
        torch.cuda.synchronize()
        return tD.cpu().numpy(), L.ftd_last_path(h).decode()
    finally:
        L.ftd_destroy(h)


def static(L, kind, A, B, C, D, P, eps_c, cyclic, stream):
    n_row, n_sys = D.shape
    tA, tB, tC, tD = dev(A), dev(B), dev(C), dev(D)
    _lib.check(L.ftd_solve_static(KINDS[kind], n_sys, n_row, P, eps_c, int(cyclic), ptr(tA),
                                  ptr(tB), ptr(tC), ptr(tD), stream))
    torch.cuda.synchronize()
    return tD.cpu().numpy()


# (kind, P, cyclic, eps_constant, (n_sys, n_row) of the first call, of the second)
CASES = [
    ("pascal", 8, False, 0.0, (4096, 512), (1000, 300)),        # streamed, then fused
    ("filtered_v2", 8, False, 1e-3, (4096, 512), (96, 256)),    # streamed (J from tile maxima)
    ("filtered", 8, False, 1e-3, (4096, 512), (2048, 256)),     # two streamed passes + x0/xN
    ("filtered", 4, True, 1e-2, (1000, 200), (333, 100)),       # fused, periodic
    ("filtered_v2", 2, False, 1e-4, (512, 300), (130, 200)),    # partitions > 64 rows: general
    ("pascal", 3, True, 0.0, (256, 600), (64, 400)),            # general PaScaL (R, E scratch)
    ("pascal", 1, True, 0.0, (128, 3000), (50, 2500)),          # general periodic Thomas (E)
]


@pytest.mark.parametrize("kind,P,cyclic,eps_c,first,second", CASES)
def test_static_equals_handle_solve(kind, P, cyclic, eps_c, first, second):
    L = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for k, (n_sys, n_row) in enumerate((first, second)):
        note(f"ftd_solve_static {kind} P={P} cyclic={cyclic} n_sys={n_sys} n_row={n_row} "
             f"call {k + 1}")
        A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, 0.05, 3.0, seed=70 + k)
        ref, path = with_handle(L, kind, A, B, C, D, P, eps_c, cyclic, stream)
        x = static(L, kind, A, B, C, D, P, eps_c, cyclic, stream)
        np.testing.assert_array_equal(x, ref, err_msg=path)


def test_static_rejects_a_solve_beyond_its_scratch():
    """Filtered v1 keeps x0/xN of every (partition, system): 2 x 16 x 600000
    doubles (154 MB) exceed the static array; refused before any launch."""
    L = _lib.load()
    n_sys, n_row, P = 600000, 512, 16
    A = torch.empty(1, dtype=torch.float64, device="cuda")  # never read: refused first
    st = L.ftd_solve_static(0, n_sys, n_row, P, 1e-3, 0, ptr(A), ptr(A), ptr(A), ptr(A), None)
    assert st == 2
    msg = L.ftd_last_error_string().decode()
    assert "bytes of scratch" in msg and str(64 << 20) in msg, msg
