"""ctypes binding of ``libzynum_gpu.so`` (the C ABI in ``include/zynum_gpu.h``).

The library is the product: there is no Python or CPU compute path behind
these functions.  Loading fails loudly when the built library is missing.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libzynum_gpu.so"

# return codes (zynum_gpu.h)
OK = 0
DECLINE_NO_DEVICE = 1
DECLINE_ALLOC = 2
DECLINE_UNSUPPORTED = 3
DECLINE_ARGUMENT = 4
DECLINE_FORCED = 5
ERR_LAUNCH = 6
DECLINE_SCRATCH = 7

F32, F64, C32, C64 = 0, 1, 2, 3
NO_TRANS, TRANS, CONJ_TRANS = 0, 1, 2
MODE_PLAIN, MODE_OZAKI_INT8, MODE_OZAKI_BF16, MODE_OZAKI_CRT, MODE_DD = 0, 1, 2, 3, 4
FLAG_ASYNC = 1

CblasRowMajor, CblasColMajor = 101, 102
CblasNoTrans, CblasTrans, CblasConjTrans = 111, 112, 113
CblasUpper, CblasLower = 121, 122
CblasNonUnit, CblasUnit = 131, 132
CblasLeft, CblasRight = 141, 142


class GemmRequest(C.Structure):
    """``zynum_gpu_gemm_request`` (104 bytes; first 88 match Zynum's GemmRequest)."""

    _fields_ = [
        ("transa", C.c_uint8), ("transb", C.c_uint8), ("scalar", C.c_uint8), ("reserved0", C.c_uint8 * 5),
        ("m", C.c_int32), ("n", C.c_int32), ("k", C.c_int32),
        ("lda", C.c_int32), ("ldb", C.c_int32), ("ldc", C.c_int32),
        ("alpha_re", C.c_uint64), ("alpha_im", C.c_uint64), ("beta_re", C.c_uint64), ("beta_im", C.c_uint64),
        ("a", C.c_void_p), ("b", C.c_void_p), ("c", C.c_void_p),
        ("stream", C.c_void_p),
        ("mode", C.c_uint8), ("mode_param", C.c_uint8), ("flags", C.c_uint8), ("tile", C.c_uint8),
        ("reserved1", C.c_uint8 * 4),
    ]


class GemmProblem(C.Structure):
    """``zynum_gpu_gemm_problem`` (72 bytes)."""

    _fields_ = [
        ("transa", C.c_uint8), ("transb", C.c_uint8), ("reserved0", C.c_uint8 * 6),
        ("m", C.c_int32), ("n", C.c_int32), ("k", C.c_int32),
        ("lda", C.c_int32), ("ldb", C.c_int32), ("ldc", C.c_int32),
        ("alpha", C.c_double), ("beta", C.c_double),
        ("a", C.c_void_p), ("b", C.c_void_p), ("c", C.c_void_p),
    ]


_I = C.POINTER(C.c_int32)
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_CH = C.c_char_p

# name -> (restype, argtypes)
SIGNATURES = {
    "cblas_sgemm": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_int,
                           C.c_void_p, C.c_int, C.c_float, C.c_void_p, C.c_int]),
    "cblas_dgemm": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int,
                           C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_int]),
    "cblas_dsyrk": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int,
                           C.c_double, C.c_void_p, C.c_int]),
    "cblas_dtrsm": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p,
                           C.c_int, C.c_void_p, C.c_int]),
    "cblas_cgemm": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                           C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "cblas_zgemm": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                           C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "cgemm_": (None, [_CH, _CH, _I, _I, _I, C.c_void_p, C.c_void_p, _I, C.c_void_p, _I, C.c_void_p, C.c_void_p, _I]),
    "zgemm_": (None, [_CH, _CH, _I, _I, _I, C.c_void_p, C.c_void_p, _I, C.c_void_p, _I, C.c_void_p, C.c_void_p, _I]),
    "cblas_ssymm": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_int, C.c_void_p,
                           C.c_int, C.c_float, C.c_void_p, C.c_int]),
    "cblas_dsymm": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p,
                           C.c_int, C.c_double, C.c_void_p, C.c_int]),
    "cblas_ssyr2k": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_int, C.c_void_p,
                            C.c_int, C.c_float, C.c_void_p, C.c_int]),
    "cblas_dsyr2k": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p,
                            C.c_int, C.c_double, C.c_void_p, C.c_int]),
    "cblas_strmm": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_void_p,
                           C.c_int, C.c_void_p, C.c_int]),
    "cblas_dtrmm": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p,
                           C.c_int, C.c_void_p, C.c_int]),
    "ssymm_": (None, [_CH, _CH, _I, _I, _F, C.c_void_p, _I, C.c_void_p, _I, _F, C.c_void_p, _I]),
    "dsymm_": (None, [_CH, _CH, _I, _I, _D, C.c_void_p, _I, C.c_void_p, _I, _D, C.c_void_p, _I]),
    "ssyr2k_": (None, [_CH, _CH, _I, _I, _F, C.c_void_p, _I, C.c_void_p, _I, _F, C.c_void_p, _I]),
    "dsyr2k_": (None, [_CH, _CH, _I, _I, _D, C.c_void_p, _I, C.c_void_p, _I, _D, C.c_void_p, _I]),
    "strmm_": (None, [_CH, _CH, _CH, _CH, _I, _I, _F, C.c_void_p, _I, C.c_void_p, _I]),
    "dtrmm_": (None, [_CH, _CH, _CH, _CH, _I, _I, _D, C.c_void_p, _I, C.c_void_p, _I]),
    "zynum_gpu_dsymm": (C.c_int, [C.c_char, C.c_char, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "zynum_gpu_dtrmm": (C.c_int, [C.c_char, C.c_char, C.c_char, C.c_char, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                  C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "zynum_gpu_dsyr2k": (C.c_int, [C.c_char, C.c_char, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p,
                                   C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "sgemm_": (None, [_CH, _CH, _I, _I, _I, _F, C.c_void_p, _I, C.c_void_p, _I, _F, C.c_void_p, _I]),
    "dgemm_": (None, [_CH, _CH, _I, _I, _I, _D, C.c_void_p, _I, C.c_void_p, _I, _D, C.c_void_p, _I]),
    "ssyrk_": (None, [_CH, _CH, _I, _I, _F, C.c_void_p, _I, _F, C.c_void_p, _I]),
    "dsyrk_": (None, [_CH, _CH, _I, _I, _D, C.c_void_p, _I, _D, C.c_void_p, _I]),
    "strsm_": (None, [_CH, _CH, _CH, _CH, _I, _I, _F, C.c_void_p, _I, C.c_void_p, _I]),
    "dtrsm_": (None, [_CH, _CH, _CH, _CH, _I, _I, _D, C.c_void_p, _I, C.c_void_p, _I]),
    "lsame_": (C.c_int, [_CH, _CH]),
    "xerbla_": (None, [_CH, _I]),
    "zynum_blas_shutdown": (None, []),
    "zynum_blas_shutdown_": (None, []),
    "zynum_gpu_gemm": (C.c_int, [C.POINTER(GemmRequest)]),
    "zynum_gpu_gemm_noalloc": (C.c_int, [C.POINTER(GemmRequest)]),
    "zynum_gpu_grouped_plan_create": (C.c_int, [C.c_int, C.POINTER(GemmProblem), C.c_int, C.c_int,
                                                C.POINTER(C.c_void_p)]),
    "zynum_gpu_grouped_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zynum_gpu_grouped_graph_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "zynum_gpu_grouped_plan_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                              C.POINTER(C.c_double)]),
    "zynum_gpu_grouped_plan_destroy": (None, [C.c_void_p]),
    "zynum_gpu_grouped_loop": (C.c_int, [C.c_int, C.POINTER(GemmProblem), C.c_int, C.c_void_p]),
    "zynum_gpu_dtrsm": (C.c_int, [C.c_char, C.c_char, C.c_char, C.c_char, C.c_int, C.c_int, C.c_double,
                                  C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "zynum_gpu_dsyrk": (C.c_int, [C.c_char, C.c_char, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int,
                                  C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "zynum_gpu_set_stream": (None, [C.c_void_p]),
    "zynum_gpu_set_copy_threads": (C.c_int, [C.c_int]),
    "zynum_gpu_last_status": (C.c_int, []),
    "zynum_gpu_debug_force_decline": (None, [C.c_int]),
    "zynum_gpu_status_string": (C.c_char_p, [C.c_int]),
    "zynum_gpu_version": (C.c_char_p, []),
    "zynum_gpu_offload_predict_ns": (C.c_double, [C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_int]),
    "zynum_gpu_offload_choose": (C.c_int, [C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_double]),
    "zynum_gpu_offload_crossover": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_double]),
    "zynum_gpu_dd_gemm": (C.c_int, [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int,
                                    C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_int, C.c_void_p]),
}

_lib = None


def load() -> C.CDLL:
    """Load the built library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def status_string(code: int) -> str:
    return load().zynum_gpu_status_string(code).decode()


class ZynumGpuError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        super().__init__(f"{what}: {status_string(code)} (code {code})" if what else status_string(code))


def check(code: int, what: str = "") -> None:
    if code != OK:
        raise ZynumGpuError(code, what)
