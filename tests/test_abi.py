"""The C ABI library without a GPU: it loads, exports every symbol that
include/ftdma.h declares, and its host-only entry points behave like the
reference's (parse table, error reporting). No kernel runs here."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from filtered_tdma_amd import _lib
from filtered_tdma_amd.backend import Kind, TdmaBackendGPU

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ftdma.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ftd_[a-z_A-Z0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libftdma.so is not built: python -m filtered_tdma_amd.build")
    return _lib.load()


def test_library_exports_every_declared_symbol(lib):
    decl = declared_symbols()
    assert len(decl) >= 15
    missing = [s for s in decl if not hasattr(lib, s)]
    assert missing == []
    assert sorted(_lib.EXPORTS) == decl


def test_exports_are_c_linkage():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    for s in declared_symbols():
        assert s in syms, s


@pytest.mark.parametrize("s,kind", [
    ("pascal", Kind.PASCAL), ("PaScaL", Kind.PASCAL), ("pascal_tdma", Kind.PASCAL),
    ("filtered_v2", Kind.FILTERED_V2), ("FILTERED_V2", Kind.FILTERED_V2),
    ("filtered", Kind.FILTERED), ("filtered_v1", Kind.FILTERED), ("filtered_tdma", Kind.FILTERED),
    ("typo", Kind.FILTERED), ("", Kind.FILTERED), (None, Kind.FILTERED),
])
def test_parse_table(lib, s, kind):
    assert TdmaBackendGPU.parse(s) == kind


def test_invalid_arguments_report_errors(lib):
    h = ctypes.c_void_p()
    assert lib.ftd_create(None, 0, 1, 1, 1, 0.0, None) == 1
    assert lib.ftd_create(ctypes.byref(h), 7, 10, 10, 1, 0.0, None) == 1
    assert b"kind" in lib.ftd_last_error_string()
    assert lib.ftd_create(ctypes.byref(h), 0, 10, 4, 8, 0.0, None) == 1
    assert b"partitions" in lib.ftd_last_error_string()
    assert lib.ftd_solve(None, None, None, None, None) == 1
    assert lib.ftd_destroy(None) == 0
    assert lib.ftd_last_comm_ms(None) == 0.0


def test_cpp_wrapper_compiles():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = ('#include "tdma_backend_gpu.hpp"\n'
           'int main() { auto k = TdmaBackendGPU::parse("pascal");'
           ' return k == TdmaBackendGPU::Kind::PASCAL ? 0 : 1; }\n')
    res = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-x", "c++", "-I",
                          os.path.join(ROOT, "include"), "-"], input=src, capture_output=True,
                         text=True)
    assert res.returncode == 0, res.stderr


def test_static_solve_rejects_oversize_before_any_device_call(lib):
    """ftd_solve_static sizes its scratch on the host and refuses a solve that
    exceeds the static array before touching the device (the pointers below
    are never dereferenced)."""
    fake = ctypes.c_void_p(1 << 20)
    st = lib.ftd_solve_static(0, 600000, 512, 16, 1e-3, 0, fake, fake, fake, fake, None)
    assert st == 2
    msg = lib.ftd_last_error_string().decode()
    assert "bytes of scratch" in msg and str(64 << 20) in msg, msg
    assert lib.ftd_solve_static(7, 10, 10, 1, 0.0, 0, fake, fake, fake, fake, None) == 1
    assert lib.ftd_solve_static(2, 10, 10, 1, 0.0, 0, None, fake, fake, fake, None) == 1
