"""Host-side mirror of Zynum's typed GEMM view API on the GPU tier.

Restates ``matrixMultiply`` / ``matrixMultiplyWithWorkspace`` /
``matrixMultiplyWorkspaceLength`` (src/blas/api/operations.zig:229,501-559)
with the checked column-major views of src/blas/api/views.zig:151-224 and the
error set of views.zig:22-30.  Storage is a flat torch tensor (device or host)
or a numpy array; the product itself always runs in ``libzynum_gpu.so``.
"""
from __future__ import annotations

import ctypes as C
import struct
from dataclasses import dataclass, replace
from typing import Any

import numpy as np

from . import _lib

INT32_MAX = 2**31 - 1


class Error(Exception):
    """Base of the view error set (views.zig:22-30)."""


class DimensionMismatch(Error):
    pass


class InvalidLength(Error):
    pass


class InvalidStride(Error):
    pass


class InvalidLeadingDimension(Error):
    pass


class BufferTooSmall(Error):
    pass


class WorkspaceTooSmall(Error):
    pass


class AliasingNotAllowed(Error):
    pass


def _storage(values: Any) -> tuple[int, int, int, str]:
    """(address, element count, element size, kind) of a flat buffer."""
    if isinstance(values, np.ndarray):
        if values.ndim != 1 or not values.flags.c_contiguous:
            raise InvalidLength("matrix storage must be a contiguous 1-D buffer")
        return values.ctypes.data, values.size, values.itemsize, values.dtype.name
    # torch tensor
    if values.dim() != 1 or not values.is_contiguous():
        raise InvalidLength("matrix storage must be a contiguous 1-D buffer")
    name = str(values.dtype).replace("torch.", "")
    return values.data_ptr(), values.numel(), values.element_size(), name


def required_matrix_storage_length(rows: int, cols: int, ld: int) -> int:
    """views.zig:79-84."""
    if rows < 0 or cols < 0:
        raise InvalidLength("negative dimension")
    if rows == 0 or cols == 0:
        return 0
    if ld < max(1, rows):
        raise InvalidLeadingDimension(f"leading dimension {ld} < {max(1, rows)}")
    return rows + (cols - 1) * ld


@dataclass(frozen=True)
class ConstMatrix:
    """Immutable column-major view (views.zig:151-192)."""

    values: Any
    row_count: int
    column_count: int
    leading_dimension: int
    operation: str = "normal"  # "normal" | "transposed" | "adjoint"

    def transposed(self) -> "ConstMatrix":
        return replace(self, operation="transposed")

    def adjoint(self) -> "ConstMatrix":
        return replace(self, operation="adjoint")

    def effective_row_count(self) -> int:
        return self.row_count if self.operation == "normal" else self.column_count

    def effective_column_count(self) -> int:
        return self.column_count if self.operation == "normal" else self.row_count

    def check(self) -> None:
        need = required_matrix_storage_length(self.row_count, self.column_count, self.leading_dimension)
        if _storage(self.values)[1] < need:
            raise BufferTooSmall("matrix storage shorter than its view")


@dataclass(frozen=True)
class Matrix:
    """Mutable column-major view (views.zig:194-224)."""

    values: Any
    row_count: int
    column_count: int
    leading_dimension: int

    def as_const(self) -> ConstMatrix:
        return ConstMatrix(self.values, self.row_count, self.column_count, self.leading_dimension)

    def check(self) -> None:
        need = required_matrix_storage_length(self.row_count, self.column_count, self.leading_dimension)
        if _storage(self.values)[1] < need:
            raise BufferTooSmall("matrix storage shorter than its view")


def _byte_range(m) -> tuple[int, int]:
    addr, _, es, _ = _storage(m.values)
    n = required_matrix_storage_length(m.row_count, m.column_count, m.leading_dimension)
    return addr, addr + n * es


def _overlap(x, y) -> bool:
    """aliasing.zig:23-26,62-64: byte ranges of the two views intersect."""
    a0, a1 = _byte_range(x)
    b0, b1 = _byte_range(y)
    if a0 == a1 or b0 == b1:
        return False
    return a0 < b1 and b0 < a1


def _check_shapes(left: ConstMatrix, right: ConstMatrix, result: Matrix) -> None:
    """operations.zig:139-143."""
    if left.effective_column_count() != right.effective_row_count():
        raise DimensionMismatch("inner dimensions differ")
    if result.row_count != left.effective_row_count():
        raise DimensionMismatch("result rows differ")
    if result.column_count != right.effective_column_count():
        raise DimensionMismatch("result columns differ")


_SCALARS = {"float64": _lib.F64, "float32": _lib.F32, "complex128": _lib.C64, "complex64": _lib.C32}


def _scalar_kind(kind: str) -> int:
    if kind in _SCALARS:
        return _SCALARS[kind]
    raise Error(f"unsupported scalar type {kind}")


def _bits(value: float, scalar: int) -> int:
    """Bit pattern of a real scalar (double, or float in the low 32 bits)."""
    if scalar in (_lib.F64, _lib.C64):
        return struct.unpack("<Q", struct.pack("<d", float(value)))[0]
    return struct.unpack("<I", struct.pack("<f", float(value)))[0]


def _trans_code(op) -> int:
    """False / True / "N" / "T" / "C" (or a view operation name) to the request code."""
    if op in (False, "N", "normal"):
        return _lib.NO_TRANS
    if op in (True, "T", "transposed"):
        return _lib.TRANS
    if op in ("C", "adjoint"):
        return _lib.CONJ_TRANS
    raise Error(f"unknown operation {op!r}")


def gemm_request(transa, transb, m: int, n: int, k: int, alpha, a: int, lda: int,
                 b: int, ldb: int, beta, c: int, ldc: int, scalar: int, stream: int | None = None,
                 mode: int = _lib.MODE_PLAIN, mode_param: int = 0, flags: int = 0, tile: int = 0) -> int:
    """Fill and submit a ``zynum_gpu_gemm_request``; returns its status code.

    ``transa`` / ``transb``: bool or "N" / "T" / "C".  ``alpha`` / ``beta`` may
    be complex for the complex scalar codes (C32 / C64)."""
    r = _lib.GemmRequest()
    r.transa = _trans_code(transa)
    r.transb = _trans_code(transb)
    r.scalar = scalar
    r.m, r.n, r.k, r.lda, r.ldb, r.ldc = m, n, k, lda, ldb, ldc
    alpha, beta = complex(alpha), complex(beta)
    r.alpha_re, r.alpha_im = _bits(alpha.real, scalar), _bits(alpha.imag, scalar)
    r.beta_re, r.beta_im = _bits(beta.real, scalar), _bits(beta.imag, scalar)
    r.a, r.b, r.c = a, b, c
    r.stream = stream
    r.mode, r.mode_param, r.flags, r.tile = mode, mode_param, flags, tile
    return _lib.load().zynum_gpu_gemm(C.byref(r))


def _run(left: ConstMatrix, right: ConstMatrix, out_addr: int, out_ld: int, rows: int, cols: int,
         product_scale, result_scale, scalar: int, stream: int | None, mode: int,
         mode_param: int) -> None:
    la, _, _, _ = _storage(left.values)
    ra, _, _, _ = _storage(right.values)
    st = gemm_request(left.operation, right.operation, rows, cols,
                      left.effective_column_count(), product_scale, la, max(1, left.leading_dimension), ra,
                      max(1, right.leading_dimension), result_scale, out_addr, max(1, out_ld), scalar, stream,
                      mode, mode_param)
    _lib.check(st, "matrix_multiply")


def matrix_multiply(left_matrix: ConstMatrix, right_matrix: ConstMatrix, result_matrix: Matrix,
                    product_scale: float | complex = 1.0, result_scale: float | complex = 0.0, *,
                    stream: int | None = None,
                    mode: int = _lib.MODE_PLAIN, mode_param: int = 0) -> None:
    """result = product_scale * op(left) * op(right) + result_scale * result (operations.zig:501-522)."""
    kinds = {_storage(x.values)[3] for x in (left_matrix, right_matrix, result_matrix)}
    if len(kinds) != 1:
        raise Error("operands must share one scalar type")
    scalar = _scalar_kind(kinds.pop())
    left_matrix.check()
    right_matrix.check()
    result_matrix.check()
    _check_shapes(left_matrix, right_matrix, result_matrix)
    if _overlap(result_matrix, left_matrix) or _overlap(result_matrix, right_matrix):
        raise AliasingNotAllowed("result overlaps an input")
    out, _, _, _ = _storage(result_matrix.values)
    _run(left_matrix, right_matrix, out, result_matrix.leading_dimension, result_matrix.row_count,
         result_matrix.column_count, product_scale, result_scale, scalar, stream, mode, mode_param)


def matrix_multiply_workspace_length(result_matrix: Matrix) -> int:
    """operations.zig:229: rows * cols of the result."""
    if result_matrix.row_count < 0 or result_matrix.column_count < 0:
        raise InvalidLength("negative dimension")
    return result_matrix.row_count * result_matrix.column_count


def matrix_multiply_with_workspace(left_matrix: ConstMatrix, right_matrix: ConstMatrix, result_matrix: Matrix,
                                   workspace: Any, product_scale: float = 1.0, result_scale: float = 0.0, *,
                                   stream: int | None = None) -> None:
    """Aliasing-tolerant form (operations.zig:529-559): C is copied to the
    workspace, the product is formed there, then copied back."""
    kinds = {_storage(x.values)[3] for x in (left_matrix, right_matrix, result_matrix)}
    if len(kinds) != 1:
        raise Error("operands must share one scalar type")
    scalar = _scalar_kind(kinds.pop())
    left_matrix.check()
    right_matrix.check()
    result_matrix.check()
    _check_shapes(left_matrix, right_matrix, result_matrix)
    need = matrix_multiply_workspace_length(result_matrix)
    waddr, wlen, _, wkind = _storage(workspace)
    if wlen < need:
        raise WorkspaceTooSmall(f"workspace has {wlen} elements, needs {need}")
    if wkind != _storage(result_matrix.values)[3]:
        raise Error("workspace scalar type differs")
    m, n = result_matrix.row_count, result_matrix.column_count
    if need == 0:
        return
    wview = Matrix(workspace[:need], m, n, m)
    for other in (left_matrix, right_matrix, result_matrix):
        if _overlap(wview, other):
            raise AliasingNotAllowed("workspace overlaps an operand")
    _copy_matrix(result_matrix.values, result_matrix.leading_dimension, workspace, m, m, n, result_scale != 0)
    _run(left_matrix, right_matrix, waddr, m, m, n, product_scale, result_scale, scalar, stream,
         _lib.MODE_PLAIN, 0)
    _copy_matrix(workspace, m, result_matrix.values, result_matrix.leading_dimension, m, n, True)


def _copy_matrix(src: Any, lds: int, dst: Any, ldd: int, m: int, n: int, enabled: bool) -> None:
    if not enabled:
        return
    for j in range(n):
        dst[j * ldd: j * ldd + m] = src[j * lds: j * lds + m]
