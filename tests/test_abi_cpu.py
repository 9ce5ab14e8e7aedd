"""The C-ABI boundary without compute: exports, request layout, xerbla text,
CBLAS enum handling and decline-before-mutation.  These paths return before
any HIP work, so they run on a host without a GPU.
"""
import ctypes as C
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "zynum_gpu.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof")))


def test_every_declared_symbol_is_exported(zlib):
    names = declared_functions()
    assert "dgemm_" in names and "cblas_dgemm" in names and "zynum_gpu_gemm" in names
    missing = [n for n in names if not hasattr(zlib, n)]
    assert missing == []


def test_exports_are_exactly_the_header(zlib):
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "zynum_amd" / "libzynum_gpu.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    exported = {e for e in exported if not e.startswith("_")}
    assert exported == set(declared_functions())


def test_request_layout():
    """GemmRequest prefix (x86_64_structured_abi.zig:70-81) plus GPU fields."""
    from zynum_amd._lib import GemmProblem, GemmRequest

    assert C.sizeof(GemmRequest) == 104
    assert GemmRequest.m.offset == 8
    assert GemmRequest.alpha_re.offset == 32
    assert GemmRequest.a.offset == 64
    assert GemmRequest.stream.offset == 88
    assert GemmRequest.mode.offset == 96
    assert C.sizeof(GemmProblem) == 72
    # the C compiler agrees with ctypes
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "zynum_gpu.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(zynum_gpu_gemm_request), offsetof(zynum_gpu_gemm_request, m),
         offsetof(zynum_gpu_gemm_request, alpha_re), offsetof(zynum_gpu_gemm_request, a),
         offsetof(zynum_gpu_gemm_request, stream), offsetof(zynum_gpu_gemm_request, mode),
         sizeof(zynum_gpu_gemm_problem));
  return 0;
}
'''
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        p = Path(d) / "l.c"
        p.write_text(src)
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(p), "-o", str(Path(d) / "l")], check=True)
        out = subprocess.run([str(Path(d) / "l")], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["104", "8", "32", "64", "88", "96", "72"]


def _child(code: str) -> subprocess.CompletedProcess:
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=str(ROOT),
                          env={"PYTHONPATH": str(ROOT), "HIP_VISIBLE_DEVICES": "", "PATH": "/usr/bin:/bin"})


@pytest.mark.parametrize("args,expected", [
    (("X", "N", 1, 1, 1, 1, 1, 1), 1), (("N", "Q", 1, 1, 1, 1, 1, 1), 2), (("N", "N", -1, 1, 1, 1, 1, 1), 3),
    (("N", "N", 1, -1, 1, 1, 1, 1), 4), (("N", "N", 1, 1, -1, 1, 1, 1), 5), (("N", "N", 3, 1, 1, 2, 1, 3), 8),
    (("T", "N", 1, 1, 3, 2, 3, 1), 8), (("N", "N", 1, 1, 3, 1, 2, 1), 10), (("N", "T", 1, 3, 1, 1, 2, 1), 10),
    (("N", "N", 4, 1, 1, 4, 1, 3), 13),
])
def test_dgemm_parameter_numbers_and_untouched_c(args, expected):
    """fortran.zig:88-103: parameter numbers; xerbla text (fortran.zig:35-39);
    C left unchanged (fortran_compat_test.zig:1103-1119)."""
    ta, tb, m, n, k, lda, ldb, ldc = args
    code = f"""
import ctypes as C, numpy as np
from zynum_amd import _lib
lib = _lib.load()
I = C.c_int32
a = np.full(16, 2.0); b = np.full(16, 3.0); c = np.full(16, 7.0)
lib.dgemm_(b"{ta}", b"{tb}", C.byref(I({m})), C.byref(I({n})), C.byref(I({k})), C.byref(C.c_double(1.0)),
           a.ctypes.data, C.byref(I({lda})), b.ctypes.data, C.byref(I({ldb})), C.byref(C.c_double(0.0)),
           c.ctypes.data, C.byref(I({ldc})))
assert (c == 7.0).all(), c
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr
    assert f" ** On entry to DGEMM parameter number {expected} had an illegal value" in r.stderr


def test_xerbla_name_trimming():
    code = """
import ctypes as C
from zynum_amd import _lib
lib = _lib.load()
lib.xerbla_(b"DTRSM ", C.byref(C.c_int32(9)))
lib.xerbla_(b"SGEMM\\x00X", C.byref(C.c_int32(3)))
lib.xerbla_(b"      ", C.byref(C.c_int32(1)))
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr
    lines = [ln for ln in r.stderr.splitlines() if ln.strip()]
    assert lines == [" ** On entry to DTRSM parameter number 9 had an illegal value",
                     " ** On entry to SGEMM parameter number 3 had an illegal value",
                     " ** On entry to UNKNOWN parameter number 1 had an illegal value"]


def test_lsame(zlib):
    assert zlib.lsame_(b"n", b"N") == 1
    assert zlib.lsame_(b"T", b"N") == 0


def test_cblas_invalid_enums_leave_c_unchanged(zlib):
    # cblas_compat_test.zig:48-61 (returns silently, no xerbla)
    a = np.array([1, 2, 3, 4.0])
    b = np.array([5, 6, 7, 8.0])
    c = np.array([9, 10, 11, 12.0])
    zlib.cblas_dgemm(999, 111, 111, 2, 2, 2, 1.0, a.ctypes.data, 2, b.ctypes.data, 2, 0.0, c.ctypes.data, 2)
    zlib.cblas_dgemm(102, 999, 111, 2, 2, 2, 1.0, a.ctypes.data, 2, b.ctypes.data, 2, 0.0, c.ctypes.data, 2)
    zlib.cblas_dtrsm(102, 141, 121, 999, 131, 2, 2, 1.0, a.ctypes.data, 2, c.ctypes.data, 2)
    assert c.tolist() == [9, 10, 11, 12]


def test_level3_invalid_parameters_leave_outputs_unchanged():
    # fortran_compat_test.zig:1184-1208 (dsyrk with trans 'X') and dtrsm side 'X'
    code = """
import ctypes as C, numpy as np
from zynum_amd import _lib
lib = _lib.load()
I = C.c_int32
a = np.array([1, 2, 3, 4.0]); c = np.array([13, 14, 15, 16.0])
lib.dsyrk_(b"U", b"X", C.byref(I(2)), C.byref(I(2)), C.byref(C.c_double(1)), a.ctypes.data, C.byref(I(2)),
           C.byref(C.c_double(0)), c.ctypes.data, C.byref(I(2)))
assert c.tolist() == [13, 14, 15, 16]
b = np.array([4, 5, 6, 7.0])
lib.dtrsm_(b"X", b"U", b"N", b"N", C.byref(I(2)), C.byref(I(2)), C.byref(C.c_double(1)), a.ctypes.data,
           C.byref(I(2)), b.ctypes.data, C.byref(I(2)))
assert b.tolist() == [4, 5, 6, 7]
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr
    assert "On entry to DSYRK parameter number 2" in r.stderr
    assert "On entry to DTRSM parameter number 1" in r.stderr


def test_decline_without_device_leaves_c_unchanged():
    """No usable device: the GPU tier declines before any write (the reference's
    'whole-operation fallback before caller-visible mutation',
    docs/architecture.md:58-60); the status is readable."""
    code = """
import ctypes as C, numpy as np
from zynum_amd import _lib, blas
lib = _lib.load()
a = np.ones(4); b = np.ones(4); c = np.full(4, 7.0)
st = blas.gemm_request(False, False, 2, 2, 2, 1.0, a.ctypes.data, 2, b.ctypes.data, 2, 0.0, c.ctypes.data, 2, _lib.F64)
assert st == _lib.DECLINE_NO_DEVICE, st
assert (c == 7.0).all()
I = C.c_int32
lib.dgemm_(b"N", b"N", C.byref(I(2)), C.byref(I(2)), C.byref(I(2)), C.byref(C.c_double(1)), a.ctypes.data,
           C.byref(I(2)), b.ctypes.data, C.byref(I(2)), C.byref(C.c_double(0)), c.ctypes.data, C.byref(I(2)))
assert lib.zynum_gpu_last_status() == _lib.DECLINE_NO_DEVICE
assert (c == 7.0).all()
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr


def test_forced_decline_hook():
    code = """
import numpy as np
from zynum_amd import _lib, blas
lib = _lib.load()
lib.zynum_gpu_debug_force_decline(1)
a = np.ones(4); c = np.full(4, 7.0)
st = blas.gemm_request(False, False, 2, 2, 2, 1.0, a.ctypes.data, 2, a.ctypes.data, 2, 0.0, c.ctypes.data, 2, _lib.F64)
assert st == _lib.DECLINE_FORCED, st
assert (c == 7.0).all()
"""
    r = _child(code)
    assert r.returncode == 0, r.stderr


def test_request_argument_checks(zlib):
    from zynum_amd import _lib, blas

    c = np.full(4, 7.0)
    st = blas.gemm_request(False, False, 2, 2, 2, 1.0, 0, 1, 0, 2, 0.0, c.ctypes.data, 2, _lib.F64)
    assert st == _lib.DECLINE_ARGUMENT  # lda < m
    assert (c == 7.0).all()
    # quick return: m == 0 touches nothing and needs no device
    st = blas.gemm_request(False, False, 0, 2, 2, 1.0, 0, 1, 0, 2, 0.0, c.ctypes.data, 1, _lib.F64)
    assert st == _lib.OK


def test_typed_view_errors():
    """views.zig:22-30 error set; operations.zig:139-143 shape checks;
    aliasing.zig overlap rule; public_surface_contract_test.zig:1340-1420."""
    from zynum_amd import blas

    buf = np.zeros(16)
    A = blas.ConstMatrix(buf[:4], 2, 2, 2)
    B = blas.ConstMatrix(np.zeros(6), 2, 3, 2)
    with pytest.raises(blas.DimensionMismatch):
        blas.matrix_multiply(A, B, blas.Matrix(np.zeros(4), 2, 2, 2))
    with pytest.raises(blas.InvalidLeadingDimension):
        blas.matrix_multiply(blas.ConstMatrix(buf, 2, 2, 1), A, blas.Matrix(np.zeros(4), 2, 2, 2))
    with pytest.raises(blas.BufferTooSmall):
        blas.matrix_multiply(blas.ConstMatrix(np.zeros(3), 2, 2, 2), A, blas.Matrix(np.zeros(4), 2, 2, 2))
    with pytest.raises(blas.AliasingNotAllowed):
        blas.matrix_multiply(A, A, blas.Matrix(buf[2:6], 2, 2, 2))
    with pytest.raises(blas.WorkspaceTooSmall):
        blas.matrix_multiply_with_workspace(A, A, blas.Matrix(np.zeros(4), 2, 2, 2), np.zeros(3))
    assert blas.matrix_multiply_workspace_length(blas.Matrix(np.zeros(12), 3, 4, 3)) == 12
    assert blas.ConstMatrix(np.zeros(6), 2, 3, 2).transposed().effective_row_count() == 3


def test_copy_threads_setter_arguments(zlib):
    """The staging copy-thread count is an API choice (no environment variable): 0 restores the
    default, 1..64 are taken, anything else declines."""
    from zynum_amd import _lib

    assert zlib.zynum_gpu_set_copy_threads(4) == _lib.OK
    assert zlib.zynum_gpu_set_copy_threads(0) == _lib.OK
    assert zlib.zynum_gpu_set_copy_threads(65) == _lib.DECLINE_ARGUMENT
    assert zlib.zynum_gpu_set_copy_threads(-1) == _lib.DECLINE_ARGUMENT
