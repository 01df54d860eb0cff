"""Host mirror of the reference's API over device-resident MPS, MPO and environments.

Names follow jdnz/MatrixProductStates.jl; Julia's mutating ``f!`` is
``f_`` here.  Every call goes through libmpship.so (HIP, gfx950); nothing in
this module computes on the host beyond building inputs.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import call, MPS_PREC_FP64, MPS_PREC_FP32

CDT = np.complex128


def _i64(seq):
    a = (C.c_int64 * len(seq))(*[int(x) for x in seq])
    return a


def _pd(arr):
    arr = np.ascontiguousarray(arr, dtype=np.float64)
    return arr, arr.ctypes.data_as(C.POINTER(C.c_double))


def _cbuf(x):
    """complex array -> column-major interleaved float64 buffer (kept alive by caller)."""
    a = np.asarray(x, dtype=CDT)
    flat = np.ascontiguousarray(a.reshape(-1, order="F"))
    v = flat.view(np.float64)
    return v, v.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """A libmpship context: one device, one HIP stream, settings."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        call("mps_ctx_create", int(device), C.byref(h))
        self.h = h
        self.device = device
        self.nb = 1

    @classmethod
    def batched(cls, nb: int, copy_bytes: int, ws_bytes: int, device: int = 0) -> "Context":
        """A lockstep batch: nb copies of every container created on it
        (copy_bytes of containers and ws_bytes of workspace per copy), every
        sweep / QR / norm call running all copies in the same launches."""
        obj = cls.__new__(cls)
        h = C.c_void_p()
        call("mps_ctx_create_batch", int(device), int(nb), int(copy_bytes), int(ws_bytes), C.byref(h))
        obj.h, obj.device, obj.nb = h, device, int(nb)
        return obj

    def companion(self) -> "Context":
        """A plain context on this context's stream, for views of single copies."""
        obj = Context.__new__(Context)
        h = C.c_void_p()
        call("mps_ctx_create_companion", self.h, C.byref(h))
        obj.h, obj.device, obj.nb = h, self.device, 1
        return obj

    def set_ragged(self, n: Sequence[int]):
        """Declare a ragged pool: copy b holds a chain of n[b] sites (n sorted
        longest first) in containers of n[0] slots, its first n[b]//2 sites
        left-aligned and the rest right-aligned (mps_ctx_set_ragged)."""
        n = [int(x) for x in n]
        if len(n) != self.nb:
            raise _lib.DimensionMismatch(-1, "one chain length per copy")
        arr = (C.c_int * len(n))(*n)
        call("mps_ctx_set_ragged", self.h, n[0], arr)
        self.ragged = n

    def close(self):
        if self.h:
            call("mps_ctx_destroy", self.h)
            self.h = None

    # No __del__: device objects keep a reference to their context, and a
    # context is released only by an explicit close().

    def set_qr_rank_tol(self, tol: float):
        call("mps_ctx_set_qr_rank_tol", self.h, float(tol))

    def set_qr_single_launch(self, on):
        """QR form: True one launch per factorization where it applies, False
        a launch per panel, None the process default."""
        call("mps_ctx_set_qr_single_launch", self.h, -1 if on is None else int(bool(on)))

    def set_precision(self, prec: str):
        call("mps_ctx_set_precision", self.h, {"fp64": MPS_PREC_FP64, "fp32": MPS_PREC_FP32}[prec])

    def set_refine_sweeps(self, n: int):
        """FP64 half-sweeps refitting the last RK4 stage after every RK4 call (even n; 0 = off)."""
        call("mps_ctx_set_refine_sweeps", self.h, int(n))

    def set_graphs(self, enable: bool):
        call("mps_ctx_set_graphs", self.h, 1 if enable else 0)

    def set_stream(self, stream_ptr):
        call("mps_ctx_set_stream", self.h, C.c_void_p(stream_ptr))

    def synchronize(self):
        call("mps_ctx_synchronize", self.h)

    def set_profiling(self, enable: bool):
        """Bracket every GEMM / QR launch with HIP events (graphs off while enabled)."""
        call("mps_ctx_set_profiling", self.h, 1 if enable else 0)

    def kernel_stats(self):
        """Sum of event-timed GEMM / QR launches since the last call (then reset)."""
        out = np.zeros(8)
        call("mps_ctx_kernel_stats", self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return {"gemm_ms": out[0], "gemm_flops": out[1], "gemm_bytes": out[2], "gemm_launches": int(out[3]),
                "qr_ms": out[4], "qr_flops": out[5], "qr_launches": int(out[6]), "launches_checked": int(out[7])}



def gemm_record(ctx: Context, enable: bool):
    call("mps_ctx_gemm_record", ctx.h, 1 if enable else 0)


def gemm_recorded(ctx: Context):
    """Distinct recorded GEMM shapes: dicts with M, N, K, batch, nb_class, cmp,
    the plan in use (tm, splits), the copies and the partials workspace."""
    rows = C.c_int()
    call("mps_ctx_gemm_recorded", ctx.h, None, 0, C.byref(rows))
    n = rows.value
    buf = (C.c_int64 * max(1, 10 * n))()
    call("mps_ctx_gemm_recorded", ctx.h, buf, n, C.byref(rows))
    keys = ("M", "N", "K", "batch", "nb_class", "cmp", "tm", "splits", "nb", "ws_bytes")
    return [dict(zip(keys, [int(x) for x in buf[10 * r:10 * r + 10]])) for r in range(n)]


def gemm_time(ctx: Context, row: int, tm: int, splits: int, iters: int = 20) -> float:
    """Microseconds per launch of recorded GEMM `row` with a forced tile and split-K depth."""
    us = C.c_double()
    call("mps_ctx_gemm_time", ctx.h, int(row), int(tm), int(splits), int(iters), C.byref(us))
    return float(us.value)


def batch_workspace_bytes(nb: int, Dmax: int, d: int, wmax: int, device: int = 0) -> int:
    """Workspace per copy of a batched context for RK4 / sweep calls at these sizes."""
    out = C.c_int64()
    call("mps_batch_workspace_bytes", int(device), int(nb), int(Dmax), int(d), int(wmax), C.byref(out))
    return int(out.value)


def _view(cls, name, base, b, ctx):
    obj = cls.__new__(cls)
    obj.__dict__.update({k: v for k, v in base.__dict__.items() if k not in ("h", "ctx")})
    h = C.c_void_p()
    call(name, base.h, int(b), ctx.h, C.byref(h))
    obj.h, obj.ctx, obj._base = h, ctx, base   # the view keeps its batched container alive
    return obj


def _ragged_len(ctx, b):
    """Chain length of copy b of a ragged pool (None for other contexts)."""
    n = getattr(ctx, "ragged", None)
    return None if n is None else n[b]


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class MPS:
    """Vector{Array{TN,3}} resident on the device (constructors.jl:10)."""

    def __init__(self, N: int, d, dims, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.N = int(N)
        self.d = [int(x) for x in (d if np.ndim(d) else [d] * N)]
        dims = [int(x) for x in dims]
        if len(dims) != self.N + 1:
            raise _lib.DimensionMismatch(-1, "dims must have N+1 entries")
        h = C.c_void_p()
        call("mps_state_create", self.ctx.h, self.N, _i64(self.d), _i64(dims), C.byref(h))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                call("mps_state_destroy", self.h)
                self.h = None
        except Exception:
            pass

    def __len__(self):
        return self.N

    @property
    def dims(self):
        out = (C.c_int64 * (self.N + 1))()
        call("mps_state_dims", self.h, out)
        return [int(x) for x in out]

    def shape(self, j):
        dm = self.dims
        return (dm[j], self.d[j], dm[j + 1])

    def upload(self, sites: Sequence[np.ndarray]):
        for j, a in enumerate(sites):
            a = np.asarray(a, dtype=CDT)
            if a.shape != self.shape(j):
                raise _lib.DimensionMismatch(-1, f"site {j}: {a.shape} != {self.shape(j)}")
            keep, p = _cbuf(a)
            call("mps_state_upload_site", self.h, j, p, *a.shape)
        return self

    def upload_f32(self, sites: Sequence[np.ndarray]):
        """Upload complex64 sites (the reference GPU examples' Complex{Float32})."""
        for j, a in enumerate(sites):
            a = np.asarray(a, dtype=np.complex64)
            if a.shape != self.shape(j):
                raise _lib.DimensionMismatch(-1, f"site {j}: {a.shape} != {self.shape(j)}")
            flat = np.ascontiguousarray(a.reshape(-1, order="F")).view(np.float32)
            call("mps_state_upload_site_f32", self.h, j, flat.ctypes.data_as(C.POINTER(C.c_float)), *a.shape)
        return self

    def site_f32(self, j) -> np.ndarray:
        """Site j as complex64 (rounded on the device)."""
        sh = self.shape(j)
        buf = np.empty(int(np.prod(sh)) * 2, dtype=np.float32)
        call("mps_state_download_site_f32", self.h, j, buf.ctypes.data_as(C.POINTER(C.c_float)), *sh)
        return buf.view(np.complex64).reshape(sh, order="F")

    def site(self, j) -> np.ndarray:
        sh = self.shape(j)
        buf = np.empty(int(np.prod(sh)) * 2, dtype=np.float64)
        call("mps_state_download_site", self.h, j, buf.ctypes.data_as(C.POINTER(C.c_double)), *sh)
        return buf.view(CDT).reshape(sh, order="F")

    def to_host(self):
        return [self.site(j) for j in range(self.N)]

    def view(self, b: int, ctx: Context) -> "MPS":
        """Copy b of a batched state, on the companion context ctx (in a ragged
        pool: copy b's own chain of n[b] sites)."""
        v = _view(MPS, "mps_state_view", self, b, ctx)
        n = _ragged_len(self.ctx, b)
        if n is not None:
            v.N, v.d = n, [self.d[0]] * n
        return v

    def copy_from(self, other: "MPS"):
        """A .= copy.(B)"""
        call("mps_state_copy", self.h, other.h)
        return self

    def similar(self) -> "MPS":
        return MPS(self.N, self.d, self.dims, self.ctx)

    @classmethod
    def from_host(cls, sites, ctx=None):
        """A device MPS of host sites; complex64 sites go through the ComplexF32 upload."""
        N = len(sites)
        dims = [s.shape[0] for s in sites] + [sites[-1].shape[2]]
        m = cls(N, [s.shape[1] for s in sites], dims, ctx)
        if all(np.asarray(s).dtype == np.complex64 for s in sites):
            return m.upload_f32(sites)
        return m.upload(sites)


class MPO:
    """Vector{Array{TN,4}} resident on the device (constructors.jl:233)."""

    def __init__(self, sites: Sequence[np.ndarray], ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.N = len(sites)
        self.d = [int(s.shape[1]) for s in sites]
        w = [int(s.shape[0]) for s in sites] + [int(sites[-1].shape[2])]
        for j in range(self.N - 1):
            if sites[j].shape[2] != sites[j + 1].shape[0]:
                raise _lib.DimensionMismatch(-1, f"MPO bond {j} mismatch")
        self.w = w
        h = C.c_void_p()
        call("mps_mpo_create", self.ctx.h, self.N, _i64(self.d), _i64(w), C.byref(h))
        self.h = h
        for j, s in enumerate(sites):
            self.upload_site(j, s)

    def upload_site(self, j, s):
        if np.asarray(s).dtype == np.complex64:
            s = np.asarray(s)
            flat = np.ascontiguousarray(s.reshape(-1, order="F")).view(np.float32)
            call("mps_mpo_upload_site_f32", self.h, j, flat.ctypes.data_as(C.POINTER(C.c_float)),
                 s.shape[0], s.shape[1], s.shape[2])
            return
        s = np.asarray(s, dtype=CDT)
        keep, p = _cbuf(s)
        call("mps_mpo_upload_site", self.h, j, p, s.shape[0], s.shape[1], s.shape[2])

    def view(self, b: int, ctx: Context) -> "MPO":
        v = _view(MPO, "mps_mpo_view", self, b, ctx)
        n = _ragged_len(self.ctx, b)
        if n is not None:
            q = lambda i: i if i <= n // 2 else i + self.N - n   # bond slot of copy b's bond i
            v.N, v.d, v.w = n, [self.d[0]] * n, [self.w[q(i)] for i in range(n + 1)]
        return v

    def set_block(self, j, wl_idx, wr_idx, op):
        """O[j][wl_idx, :, wr_idx, :] = op (update_h! / update_L_Ryd!)."""
        keep, p = _cbuf(np.asarray(op, dtype=CDT))
        call("mps_mpo_set_block", self.h, j, int(wl_idx), int(wr_idx), p)

    def __del__(self):
        try:
            if self.h:
                call("mps_mpo_destroy", self.h)
                self.h = None
        except Exception:
            pass

    def __len__(self):
        return self.N


class Env:
    """build_env result resident on the device (constructors.jl:345-375)."""

    def __init__(self, dims1, dims2, mpo_dims=None, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.N = len(dims1) - 1
        self.op = mpo_dims is not None
        h = C.c_void_p()
        if self.op:
            call("mps_env_create_op", self.ctx.h, self.N, _i64(dims1), _i64(dims2),
                 _i64([int(x) for x in mpo_dims]), C.byref(h))
        else:
            call("mps_env_create", self.ctx.h, self.N, _i64(dims1), _i64(dims2), C.byref(h))
        self.h = h
        self.dims1, self.dims2 = list(dims1), list(dims2)
        self.mpo_dims = list(mpo_dims) if mpo_dims is not None else None

    def view(self, b: int, ctx: Context) -> "Env":
        v = _view(Env, "mps_env_view", self, b, ctx)
        n = _ragged_len(self.ctx, b)
        if n is not None:
            q = lambda i: i if i <= n // 2 else i + self.N - n
            v.N = n
            v.dims1 = [self.dims1[q(i)] for i in range(n + 1)]
            v.dims2 = [self.dims2[q(i)] for i in range(n + 1)]
            if self.mpo_dims is not None:
                v.mpo_dims = [self.mpo_dims[q(i)] for i in range(n + 1)]
        return v

    def slot(self, j, shape):
        n = int(np.prod(shape))
        buf = np.empty(2 * n, dtype=np.float64)
        call("mps_env_download_slot", self.h, j, buf.ctypes.data_as(C.POINTER(C.c_double)), n)
        return buf.view(CDT).reshape(shape, order="F")

    def __del__(self):
        try:
            if self.h:
                call("mps_env_destroy", self.h)
                self.h = None
        except Exception:
            pass


# --------------------------------------------------------------- constructors

def _dims_rule(na, dmax, d):
    dims = np.ones(na + 2, dtype=np.int64)
    dims[2:na + 1] = dmax
    for n in range(1, na):
        dims[n + 1] = min(dims[n] * d, dims[n + 1])
        dims[na + 1 - n] = min(dims[na + 2 - n] * d, dims[na + 1 - n])
    return [int(x) for x in dims[1:]]


def mpsproductstate(na, dmax, d, state, ctx=None):
    """constructors.jl:5-61: product state, left canonical; returns (mps, dims)."""
    dims = _dims_rule(na, dmax, d)
    m = MPS(na, d, dims, ctx)
    vecs = np.zeros((na, d), dtype=CDT)
    for n in range(na):
        vecs[n] = state(n + 1) if callable(state) else state
    keep, p = _cbuf(vecs.T)   # states[n*d + s]
    call("mps_state_product", m.h, p)
    return m, dims


def mpsgroundstate(na, dmax, d, ctx=None):
    """constructors.jl:67-75."""
    st = np.zeros(d, dtype=CDT)
    st[0] = 1.0
    return mpsproductstate(na, dmax, d, st, ctx)


def mpsrandom(na, dmax, d, rng, ctx=None):
    """constructors.jl:111-129: random sites (host RNG), left-orthonormalised on device."""
    dims = np.ones(na + 2, dtype=np.int64)
    dims[2:na + 1] = dmax
    shapes = []
    for n in range(1, na + 1):
        dims[n + 1] = min(dims[n] * d, dims[n + 1])
        dims[na + 1 - n] = min(dims[na + 2 - n] * d, dims[na + 1 - n])
        shapes.append((int(dims[n]), d, int(dims[n + 1])))
    dl = [int(x) for x in dims[1:]]
    m = MPS(na, d, dl, ctx)
    s = np.sqrt(0.5)
    sites = [(s * rng.standard_normal(sh) + 1j * s * rng.standard_normal(sh)).astype(CDT) for sh in shapes]
    m.upload(sites)
    for j in range(na):
        call("mps_leftorth_site", m.h, j)
    return m, dl


def makempo(mpomat, na, d, ctx=None):
    """constructors.jl:229-253 (host assembly, device residence)."""
    b1, b2 = len(mpomat), len(mpomat[0])
    sites = []
    for n in range(1, na + 1):
        s = np.zeros((b1, d, b2, d), dtype=CDT)
        for mm in range(b2):
            for ll in range(b1):
                s[ll, :, mm, :] = mpomat[ll][mm](n)
        sites.append(s)
    sites[0] = sites[0][0:1].copy()
    sites[-1] = sites[-1][:, :, b2 - 1:b2, :].copy()
    return MPO(sites, ctx)


def build_env(dims1, dims2, mpo_dims=None, ctx=None):
    """constructors.jl:345-375."""
    return Env(dims1, dims2, mpo_dims, ctx)


def mpsdims(mps: MPS):
    """MatrixProductStates.jl:41-51."""
    return mps.dims


# ------------------------------------------------------------ gauge moves

def leftorth_(mps: MPS, site: int):
    """orthonormalisation.jl:164-171 on site ``site`` of a device MPS."""
    call("mps_leftorth_site", mps.h, int(site))


def rightorth_(mps: MPS, site: int):
    """orthonormalisation.jl:227-234 on site ``site`` of a device MPS."""
    call("mps_rightorth_site", mps.h, int(site))


def _scalar(name, mps):
    out = C.c_double()
    call(name, mps.h, C.byref(out))
    return float(out.value)


def leftorthogonalizeQR_(mps):
    """orthonormalisation.jl:43-56."""
    return _scalar("mps_leftorthogonalize_qr", mps)


def leftorthonormalizationQR_(mps):
    """orthonormalisation.jl:29-36."""
    return _scalar("mps_leftorthonormalize_qr", mps)


def rightorthogonalizeQR_(mps):
    """orthonormalisation.jl:77-90."""
    return _scalar("mps_rightorthogonalize_qr", mps)


def rightorthonormalizationQR_(mps):
    """orthonormalisation.jl:64-71."""
    return _scalar("mps_rightorthonormalize_qr", mps)


def normalize_lo_(mps):
    """orthonormalisation.jl:97-103."""
    return _scalar("mps_normalize_lo", mps)


def normalize_lo_batch(mps) -> np.ndarray:
    """normalize_lo! of every copy of a batched state; norm^2 per copy."""
    out = np.zeros(mps.ctx.nb)
    call("mps_normalize_lo_batch", mps.h, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def norm_lo(mps):
    """orthonormalisation.jl:117-121."""
    return _scalar("mps_norm_lo", mps)


# ------------------------------------------------------------------ sweeps

def _coef(coef):
    c = np.asarray(coef, dtype=CDT)
    return _pd(np.ascontiguousarray(c).view(np.float64))


def _handles(objs):
    return (C.c_void_p * len(objs))(*[o.h.value for o in objs])


def compress_var_apply_H_left_(A0, envop, Ain, A, H):
    """compression.jl:321-347."""
    call("mps_compress_var_apply_H_left", A0.h, envop.h, Ain.h, A.h, H.h)


def compress_sum_var_apply_H_left_(A0, env, envop, Ain, coef, A, H):
    """compression.jl:889-944."""
    keep, pc = _coef(coef)
    call("mps_compress_sum_var_apply_H_left", A0.h, _handles(env), envop.h, Ain.h, len(coef), pc,
         _handles(A), H.h)


def compress_sum_var_apply_H_right_(A0, env, envop, Ain, coef, A, H):
    """compression.jl:1007-1063."""
    keep, pc = _coef(coef)
    call("mps_compress_sum_var_apply_H_right", A0.h, _handles(env), envop.h, Ain.h, len(coef), pc,
         _handles(A), H.h)


def compress_var_apply_H_(A0, envop, Ain, A, H, maxsweeps):
    """compression.jl:217-258."""
    call("mps_compress_var_apply_H", A0.h, envop.h, Ain.h, A.h, H.h, int(maxsweeps))


def compress_var_(A0, A, env, maxsweeps):
    """compression.jl:65-103."""
    call("mps_compress_var", A0.h, A.h, env.h, int(maxsweeps))


def compress_sum_var(Ain, coef, A, maxsweeps):
    """compression.jl:496-565 (returns a new device MPS)."""
    out = Ain.similar()
    keep, pc = _coef(coef)
    call("mps_compress_sum_var", out.h, Ain.h, len(coef), pc, _handles(A), int(maxsweeps))
    return out


def compress_sum_var_apply_H_(A0, envop, coef, A, H, maxsweeps):
    """compression.jl:805-887 (by intent)."""
    keep, pc = _coef(coef)
    call("mps_compress_sum_var_apply_H", A0.h, envop.h, len(coef), pc, _handles(A), H.h, int(maxsweeps))


def compressionSVD_(mps, dmax, tol=1e-15):
    """compression.jl:1280-1312 -> (err_glob, err_trim, ent_ent)."""
    eg = C.c_double()
    et = np.zeros(mps.N)
    ee = np.zeros(mps.N)
    call("mps_compression_svd", mps.h, int(dmax), float(tol), C.byref(eg),
         et.ctypes.data_as(C.POINTER(C.c_double)), ee.ctypes.data_as(C.POINTER(C.c_double)))
    return float(eg.value), et, ee


def RK4stp_apply_H_half_(dt, H, *args):
    """timeevolution.jl:178-196.

    10-argument form ``(dt, H, psi, psi1, psi2, psi3, env1, env2, env3, envop)``
    overwrites ``psi``; 11-argument form ``(dt, H, psi0, psi, ...)`` writes ``psi0``.
    """
    if len(args) == 8:
        psi0 = None
        psi, psi1, psi2, psi3, env1, env2, env3, envop = args
    elif len(args) == 9:
        psi0, psi, psi1, psi2, psi3, env1, env2, env3, envop = args
    else:
        raise TypeError("RK4stp_apply_H_half!: 10 or 11 arguments")
    H1, H2, H3 = H
    call("mps_rk4stp_apply_H_half", float(dt), H1.h, H2.h, H3.h, psi0.h if psi0 is not None else None,
         psi.h, psi1.h, psi2.h, psi3.h, env1.h, env2.h, env3.h, envop.h)


def Arnoldi_expm_(A, rhov, numv, env, envop):
    """timeevolution.jl:88-119; returns the Krylov coefficients."""
    coef = np.zeros(2 * numv)
    call("mps_arnoldi_expm", A.h, int(numv), _handles(rhov[:numv]), env.h if env is not None else None,
         envop.h, coef.ctypes.data_as(C.POINTER(C.c_double)))
    return coef.view(CDT)


# ---------------------------------------------------- two-site sweeps
def _two_site(Ain, coef, A, H, mode, maxsweeps, svd_tol, dims, dlim):
    """A new device MPS from mps_compress_two_site (capacity: the dims rule
    min(d^n, dlim, d^(N-n)) and Ain's bonds); the reference's ``dims`` list
    is updated in place to the new bonds, as the reference does."""
    N, d, ctx = Ain.N, Ain.d, Ain.ctx
    cap = [1] * (N + 1)
    left = right = 1
    for n in range(1, N):
        left *= d[n - 1]
        cap[n] = min(left, int(dlim))
    for n in range(N - 1, 0, -1):
        right *= d[n]
        cap[n] = min(cap[n], right)
    cap = [max(c, a) for c, a in zip(cap, Ain.dims)]
    A0 = MPS(N, d, cap, ctx)
    nov = len(A) - (1 if H is not None else 0)
    envs = [build_env(cap, A[l].dims, ctx=ctx) for l in range(nov)]
    envop = build_env(cap, A[nov].dims, H.w, ctx=ctx) if H is not None else None
    keep, pc = _coef(coef)
    call("mps_compress_two_site", A0.h, _handles(envs) if envs else None, envop.h if envop else None, Ain.h,
         len(A), pc, _handles(A), H.h if H is not None else None, int(mode), int(maxsweeps), float(svd_tol),
         int(dlim))
    if dims is not None:
        dims[:] = A0.dims
    return A0


def compress_var_two_site(Ain, A, maxsweeps, svd_tol, dims, dlim):
    """compression.jl:110-164 (by intent: thin SVD for svdecon)."""
    return _two_site(Ain, [1.0], [A], None, 0, maxsweeps, svd_tol, dims, dlim)


def compress_var_apply_H_two_site(Ain, A, H, maxsweeps, svd_tol, dims, dlim):
    """compression.jl:355-411."""
    return _two_site(Ain, [1.0], [A], H, 0, maxsweeps, svd_tol, dims, dlim)


def compress_var_apply_H_two_site_left(Ain, A, H, maxsweeps, svd_tol, dims, dlim):
    """compression.jl:413-448 (one left half-sweep)."""
    return _two_site(Ain, [1.0], [A], H, 1, 1, svd_tol, dims, dlim)


def compress_var_apply_H_two_site_right(Ain, A, H, maxsweeps, svd_tol, dims, dlim):
    """compression.jl:450-486 (one right half-sweep; Ain right canonical)."""
    return _two_site(Ain, [1.0], [A], H, 2, 1, svd_tol, dims, dlim)


def compress_sum_var_two_site(Ain, coef, A, maxsweeps, svd_tol, dims, dlim):
    """compression.jl:574-647."""
    return _two_site(Ain, coef, A, None, 0, maxsweeps, svd_tol, dims, dlim)


def compress_sum_var_apply_H_two_site(Ain, coef, A, H, maxsweeps, svd_tol, dims, dlim):
    """compression.jl:1072-1157."""
    return _two_site(Ain, coef, A, H, 0, maxsweeps, svd_tol, dims, dlim)


def compress_sum_var_apply_H_two_site_left(Ain, coef, A, H, maxsweeps, svd_tol, dims, dlim):
    """compression.jl:1159-1216."""
    return _two_site(Ain, coef, A, H, 1, 1, svd_tol, dims, dlim)


def compress_sum_var_apply_H_two_site_right(Ain, coef, A, H, maxsweeps, svd_tol, dims, dlim):
    """compression.jl:1218-1273 (Ain right canonical)."""
    return _two_site(Ain, coef, A, H, 2, 1, svd_tol, dims, dlim)


def RK4stp_apply_H_half_full_two_site(dt, H, psi, svd_tol, dims, dlim):
    """timeevolution.jl:208-216: the RK4 half step with two-site sweeps in
    every stage; returns the new state."""
    psi1 = compress_var_apply_H_two_site_left(psi, psi, H[0], 1, svd_tol, dims, dlim)
    psi2 = compress_sum_var_apply_H_two_site_right(psi1, [1.0, dt / 2], [psi, psi1], H[1], 1, svd_tol, dims, dlim)
    psi3 = compress_sum_var_apply_H_two_site_left(psi2, [1.0, dt], [psi, psi2], H[1], 1, svd_tol, dims, dlim)
    return compress_sum_var_apply_H_two_site_right(psi3, [-1.0 / 3, 1.0 / 3, 2.0 / 3, 1.0 / 3],
                                                   [psi, psi1, psi2, psi3], H[2], 1, svd_tol, dims, dlim)


def steady_state_search_(At, L, Lconj, maxsweeps, converge_rat, eigs_tol, envop=None):
    """steadystate.jl:6-86.  ``L`` and ``Lconj`` are host MPO site lists (the
    reference passes its MPO arrays); their product W = Lconj * L is formed on
    the host (applyMPOtoMPO) and the search runs on the device state ``At``.
    Returns (last local eigenvalue, sweeps run)."""
    from .observables import applyMPOtoMPO
    W = MPO(applyMPOtoMPO(L, Lconj), At.ctx)
    if envop is None:
        envop = build_env(At.dims, At.dims, W.w, ctx=At.ctx)
    ev = C.c_double()
    sw = C.c_int()
    call("mps_steady_state_search", At.h, W.h, envop.h, int(maxsweeps), float(converge_rat), float(eigs_tol),
         C.byref(ev), C.byref(sw))
    return float(ev.value), int(sw.value)


# --------------------------------------------------------- scalar products

def _c2(name, *h):
    out = np.zeros(2)
    call(name, *h, out.ctypes.data_as(C.POINTER(C.c_double)))
    return complex(out[0], out[1])


def scal_prod(mps1, mps2):
    """MatrixProductStates.jl:345-358: <mps1|mps2>."""
    return _c2("mps_scal_prod", mps1.h, mps2.h)


def scal_op_prod(mps1, op, mps2):
    """MatrixProductStates.jl:327-340: conj(<mps1|op|mps2>)."""
    return _c2("mps_scal_op_prod", mps1.h, op.h, mps2.h)


def scal_prod_no_conj(mps1, mps2):
    """MatrixProductStates.jl:366-379."""
    return _c2("mps_scal_prod_no_conj", mps1.h, mps2.h)


def normMPS(mps):
    """MatrixProductStates.jl:317-322."""
    return abs(scal_prod(mps, mps))


def diff_norm(mps1, mps2):
    """MatrixProductStates.jl:388-392."""
    return normMPS(mps1) + normMPS(mps2) - 2 * scal_prod(mps1, mps2).real


def scal_prod_no_conj_noalloc(mps1, mps2):
    """scal_prod_no_conj in the library's fixed device buffer (no device allocation)."""
    return _c2("mps_scal_prod_no_conj_noalloc", mps1.h, mps2.h)


def measure_excitations_(out, mps, op):
    """spin_model_quantum_trajectory_rk4.jl:220-231 (left-canonical, normalised mps)."""
    keep, p = _cbuf(np.asarray(op, dtype=CDT))
    res = np.zeros(mps.N)
    call("mps_measure_local", mps.h, p, res.ctypes.data_as(C.POINTER(C.c_double)))
    if out is not None:
        out[:] = res
    return res


def apply_site_operator_(mps, site, op):
    """MatrixProductStates.jl:135-143 on site ``site`` of a device MPS."""
    keep, p = _cbuf(np.asarray(op, dtype=CDT))
    call("mps_apply_site_operator", mps.h, int(site), p)
