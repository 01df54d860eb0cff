"""mpsamd -- MI355X (gfx950) hot path of jdnz/MatrixProductStates.jl.

Device-resident MPS/MPO/environments and the reference's sweep, RK4,
Arnoldi, gauge and scalar-product functions, backed by libmpship.so.
Julia's mutating ``f!`` is spelled ``f_``.
"""
from . import models  # noqa: F401  (host-side model builders, numpy only)
from ._lib import MPSError, DimensionMismatch, load  # noqa: F401
from .core import (  # noqa: F401
    Context, MPS, MPO, Env, default_context, batch_workspace_bytes, normalize_lo_batch,
    mpsproductstate, mpsgroundstate, mpsrandom, makempo, build_env, mpsdims,
    leftorth_, rightorth_, leftorthogonalizeQR_, leftorthonormalizationQR_,
    rightorthogonalizeQR_, rightorthonormalizationQR_, normalize_lo_, norm_lo,
    compress_var_apply_H_left_, compress_sum_var_apply_H_left_, compress_sum_var_apply_H_right_,
    compress_var_apply_H_, compress_var_, compress_sum_var, compress_sum_var_apply_H_,
    compressionSVD_, RK4stp_apply_H_half_, Arnoldi_expm_, steady_state_search_,
    compress_var_two_site, compress_var_apply_H_two_site, compress_var_apply_H_two_site_left,
    compress_var_apply_H_two_site_right, compress_sum_var_two_site, compress_sum_var_apply_H_two_site,
    compress_sum_var_apply_H_two_site_left, compress_sum_var_apply_H_two_site_right,
    RK4stp_apply_H_half_full_two_site,
    scal_prod, scal_op_prod, scal_prod_no_conj, normMPS, diff_norm,
    measure_excitations_, apply_site_operator_,
)
