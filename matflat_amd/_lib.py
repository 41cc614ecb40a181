"""ctypes binding of libmatflat_hip.so (declarations in include/matflat_hip.h).

The library is built in-tree by ``python -m matflat_amd.build``.  Loading it
fails loudly when it is missing: there is no CPU fallback anywhere in this
package.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmatflat_hip.so")

MFH_OK = 0
MFH_ERR_HIP = -1001
MFH_ERR_ALLOC = -1002
MFH_ERR_UNSUPPORTED = -1003
MFH_ERR_STATE = -1004

NOTRANS, TRANS, CONJTRANS, CONJNOTRANS = 111, 112, 113, 114
UPPER, LOWER = 121, 122
NONUNIT, UNIT = 131, 132
EXACT_MAX_N = 1024
SMALL_MAX_N = 128
LDS_MAX_N = 128  # larger orders run the generic kernels (k_getrf_gen, k_potrf_gen, k_trsv_gen)
MATH_EXACT, MATH_FAST = 0, 1
OP_DPOSV, OP_DGETRF, OP_DPOTRF, OP_DTRSV, OP_DGESV, OP_DGESV_MIXED, OP_DGEMM, OP_ZGEMM = 1, 2, 3, 4, 5, 6, 7, 8
NOALLOC_MAX_COUNT = 1 << 20
BLOCKED_MAX_N = 512
KERNEL_SORT, KERNEL_CHOL_EXACT, KERNEL_CHOL_FAST, KERNEL_GETRF_EXACT, KERNEL_TRSV_EXACT, KERNEL_GETRS_EXACT = \
    1, 2, 3, 4, 5, 6
KERNEL_GESV_MIXED, KERNEL_GESV_F64, KERNEL_GESV_STEP, KERNEL_GEMV_EXACT, KERNEL_GETRI_SETUP = 7, 8, 9, 10, 11
KERNEL_GETRI_EXACT, KERNEL_GETRF_GEN, KERNEL_POTRF_GEN, KERNEL_TRSV_GEN, KERNEL_GETRS_PERMUTE = 12, 13, 14, 15, 16
KERNEL_GEVD_EXACT = 17
KERNEL_NAMES = {1: "size sort", 2: "k_chol_packed", 3: "k_chol_fast", 4: "k_getrf_exact", 5: "k_trsv_exact",
                6: "k_getrs_exact", 7: "k_gesv<float, mixed>", 8: "k_gesv<double>", 9: "k_gesv_step",
                10: "k_dgemv_exact", 11: "k_getri_setup", 12: "k_getri_exact", 13: "k_getrf_gen",
                14: "k_potrf_gen", 15: "k_trsv_gen", 16: "k_getrs_permute", 17: "k_gevd_exact"}

_i, _d, _p, _i64 = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int64

# name -> argtypes (all return int)
SIGNATURES = {
    "mfh_create": [_p],
    "mfh_destroy": [_p],
    "mfh_set_stream": [_p, _p],
    "mfh_get_stream": [_p, _p],
    "mfh_synchronize": [_p],
    "mfh_version": [],
    "mfh_set_math_mode": [_p, _i],
    "mfh_get_math_mode": [_p, _p],
    "mfh_set_fast_waves": [_p, _i, _i],
    "mfh_get_fast_waves": [_p, _i, _p],
    "mfh_set_bucket_streams": [_p, _i],
    "mfh_set_chol_shape": [_p, _i, _i, _i],
    "mfh_get_bucket_streams": [_p, _p],
    "mfh_set_timing": [_p, _i],
    "mfh_get_timing": [_p, _i, _p, _p, _p, _p],
    "mfh_dgetrf_vbatched": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _i, _i],
    "mfh_dgetrf_strided_batched": [_p, _i, _i, _i, _p, _i, _i64, _p, _i64, _p, _p],
    "mfh_dpotrf_vbatched": [_p, _i, _p, _p, _p, _p, _i],
    "mfh_dtrsv_vbatched": [_p, _i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _i],
    "mfh_dposv_vbatched": [_p, _i, _p, _p, _p, _p, _p, _i],
    "mfh_dposv_vbatched_noalloc": [_p, _i, _p, _p, _p, _p, _p, _i],
    "mfh_dgetrs_vbatched": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _i],
    "mfh_dgemm_vbatched": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i],
    "mfh_zgemm_vbatched": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i],
    "mfh_dgetri_vbatched": [_p, _i, _p, _p, _p, _p, _p, _i],
    "mfh_dgemv_vbatched": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i],
    "mfh_zgetrf_vbatched": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _i, _i],
    "mfh_zpotrf_vbatched": [_p, _i, _p, _p, _p, _p, _i],
    "mfh_ztrsv_vbatched": [_p, _i, _i, _i, _p, _p, _p, _p, _p, _p, _i],
    "mfh_dgesv_mixed_vbatched": [_p, _i, _p, _p, _p, _p, _p, _i, _d, _p, _p, _p, _i],
    "mfh_dgesv_vbatched": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _i],
    "mfh_dgetrf_blocked_vbatched": [_p, _i, _p, _p, _p, _p, _p, _i],
    "mfh_dposv_backward_vbatched": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i],
    "mfh_dgevd_vbatched": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _i],
    "mfh_dpotrf_backward_vbatched": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _i],
    "mfh_plan_create_dposv": [_p, _i, _p, _p, _p, _p, _p, _i, _p],
    "mfh_plan_create_dgemm": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _p],
    "mfh_plan_create_zgemm": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _p],
    "mfh_plan_create_dgetrf": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p],
    "mfh_plan_create_dpotrf": [_p, _i, _p, _p, _p, _p, _i, _p],
    "mfh_plan_create_dtrsv": [_p, _i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _i, _p],
    "mfh_plan_create_dgesv": [_p, _i, _p, _p, _p, _p, _p, _p, _p, _i, _p],
    "mfh_plan_create_dgesv_mixed": [_p, _i, _p, _p, _p, _p, _p, _i, _d, _p, _p, _p, _i, _p],
    "mfh_plan_get_op": [_p, _p],
    "mfh_plan_execute": [_p],
    "mfh_plan_destroy": [_p],
}

_lib = None


def load() -> ctypes.CDLL:
    """Load libmatflat_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m matflat_amd.build` "
                "(the HIP library is required; there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        lib.mfh_last_error.restype = ctypes.c_char_p
        lib.mfh_last_error.argtypes = []
        _lib = lib
    return _lib


def last_error() -> str:
    return load().mfh_last_error().decode(errors="replace")
