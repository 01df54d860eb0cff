"""ctypes binding of libmpship.so (the C ABI declared in include/mpship.h).

The library is loaded from this package's ``lib/`` directory (built in-tree by
``mpsamd.build``).  There is no fallback: if the library is missing or fails
to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libmpship.so")

MPS_OK = 0
MPS_E_ARG, MPS_E_HIP, MPS_E_ALLOC, MPS_E_STATE, MPS_E_NUMERIC = -1, -2, -3, -4, -5
MPS_PREC_FP64, MPS_PREC_FP32 = 0, 1

vp = C.c_void_p
i64 = C.c_int64
i32 = C.c_int
dbl = C.c_double
P64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)
PF = C.POINTER(C.c_float)
PVP = C.POINTER(C.c_void_p)

# name -> argtypes (all return int status except mps_last_error)
SIGNATURES = {
    "mps_version": [],
    "mps_ctx_create": [i32, PVP],
    "mps_ctx_create_batch": [i32, i32, i64, i64, PVP],
    "mps_batch_workspace_bytes": [i32, i32, i64, i64, i64, P64],
    "mps_ctx_create_companion": [vp, PVP],
    "mps_ctx_set_ragged": [vp, i32, C.POINTER(C.c_int)],
    "mps_state_view": [vp, i32, vp, PVP],
    "mps_env_view": [vp, i32, vp, PVP],
    "mps_mpo_view": [vp, i32, vp, PVP],
    "mps_normalize_lo_batch": [vp, PD],
    "mps_ctx_destroy": [vp],
    "mps_ctx_set_stream": [vp, vp],
    "mps_ctx_synchronize": [vp],
    "mps_ctx_set_qr_rank_tol": [vp, dbl],
    "mps_ctx_set_qr_single_launch": [vp, i32],
    "mps_ctx_set_precision": [vp, i32],
    "mps_ctx_set_graphs": [vp, i32],
    "mps_ctx_set_refine_sweeps": [vp, i32],
    "mps_ctx_set_profiling": [vp, i32],
    "mps_ctx_kernel_stats": [vp, PD],
    "mps_ctx_gemm_record": [vp, i32],
    "mps_ctx_gemm_recorded": [vp, P64, i32, C.POINTER(C.c_int)],
    "mps_ctx_gemm_time": [vp, i32, i32, i32, i32, PD],
    "mps_state_create": [vp, i32, P64, P64, PVP],
    "mps_state_destroy": [vp],
    "mps_state_dims": [vp, P64],
    "mps_state_upload_site": [vp, i32, PD, i64, i64, i64],
    "mps_state_download_site": [vp, i32, PD, i64, i64, i64],
    "mps_state_upload_site_f32": [vp, i32, PF, i64, i64, i64],
    "mps_state_download_site_f32": [vp, i32, PF, i64, i64, i64],
    "mps_mpo_upload_site_f32": [vp, i32, PF, i64, i64, i64],
    "mps_state_copy": [vp, vp],
    "mps_state_product": [vp, PD],
    "mps_mpo_create": [vp, i32, P64, P64, PVP],
    "mps_mpo_destroy": [vp],
    "mps_mpo_upload_site": [vp, i32, PD, i64, i64, i64],
    "mps_mpo_set_block": [vp, i32, i64, i64, PD],
    "mps_env_create_op": [vp, i32, P64, P64, P64, PVP],
    "mps_env_create": [vp, i32, P64, P64, PVP],
    "mps_env_destroy": [vp],
    "mps_env_download_slot": [vp, i32, PD, i64],
    "mps_dev_prod_AOR": [vp, vp, vp, i64, i64, i64, vp, i64, i64, vp, i64],
    "mps_dev_prod_AOL": [vp, vp, vp, i64, i64, i64, vp, i64, i64, vp, i64],
    "mps_dev_prod_LR_34": [vp, vp, vp, vp, i64, i64, i64, i64, i64],
    "mps_dev_prod_LR_43": [vp, vp, vp, vp, i64, i64, i64, i64, i64],
    "mps_dev_prod_LR_23": [vp, vp, vp, vp, i64, i64, i64, i64, i32],
    "mps_dev_prod_LR_32": [vp, vp, vp, vp, i64, i64, i64, i64],
    "mps_dev_update_re": [vp, vp, vp, vp, i64, i64, i64, i64, i64],
    "mps_dev_update_le": [vp, vp, vp, vp, i64, i64, i64, i64, i64],
    "mps_dev_update_renv": [vp, vp, vp, vp, i64, i64, i64, i64],
    "mps_dev_update_lenv": [vp, vp, vp, vp, i64, i64, i64, i64, i32],
    "mps_dev_local_exp": [vp, PD, vp, vp, i64, i64, i64, PD],
    "mps_dev_qrq": [vp, vp, i64, i64, i64, i64, vp],
    "mps_leftorth_site": [vp, i32],
    "mps_rightorth_site": [vp, i32],
    "mps_leftorthogonalize_qr": [vp, PD],
    "mps_leftorthonormalize_qr": [vp, PD],
    "mps_rightorthogonalize_qr": [vp, PD],
    "mps_rightorthonormalize_qr": [vp, PD],
    "mps_normalize_lo": [vp, PD],
    "mps_norm_lo": [vp, PD],
    "mps_compress_var_apply_H_left": [vp, vp, vp, vp, vp],
    "mps_compress_sum_var_apply_H_left": [vp, PVP, vp, vp, i32, PD, PVP, vp],
    "mps_compress_sum_var_apply_H_right": [vp, PVP, vp, vp, i32, PD, PVP, vp],
    "mps_compress_var_apply_H": [vp, vp, vp, vp, vp, i32],
    "mps_compress_var": [vp, vp, vp, i32],
    "mps_compress_sum_var": [vp, vp, i32, PD, PVP, i32],
    "mps_compress_sum_var_apply_H": [vp, vp, i32, PD, PVP, vp, i32],
    "mps_compression_svd": [vp, i64, dbl, PD, PD, PD],
    "mps_rk4stp_apply_H_half": [dbl, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
    "mps_arnoldi_expm": [vp, i32, PVP, vp, vp, PD],
    "mps_compress_two_site": [vp, PVP, vp, vp, i32, PD, PVP, vp, i32, i32, dbl, i64],
    "mps_steady_state_search": [vp, vp, vp, i32, dbl, dbl, PD, C.POINTER(C.c_int)],
    "mps_scal_prod": [vp, vp, PD],
    "mps_scal_op_prod": [vp, vp, vp, PD],
    "mps_scal_prod_no_conj": [vp, vp, PD],
    "mps_scal_prod_no_conj_noalloc": [vp, vp, PD],
    "mps_measure_local": [vp, PD, PD],
    "mps_apply_site_operator": [vp, i32, PD],
}


class MPSError(RuntimeError):
    """A non-zero status from libmpship (message from mps_last_error)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class DimensionMismatch(MPSError, ValueError):
    """Status MPS_E_ARG -- the reference raises DimensionMismatch here."""


_lib = None


def load(path: str | None = None):
    """Load libmpship.so (once) and bind every declared entry point."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"libmpship.so not built ({p}); run mpsamd.build.build()")
    lib = C.CDLL(p)
    for name, argt in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = argt
        f.restype = C.c_int
    lib.mps_last_error.argtypes = []
    lib.mps_last_error.restype = C.c_char_p
    if path is None:
        _lib = lib
    return lib


def check(rc):
    if rc != MPS_OK:
        msg = load().mps_last_error().decode(errors="replace")
        if rc == MPS_E_ARG:
            raise DimensionMismatch(rc, msg)
        raise MPSError(rc, msg)
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args))
