"""L1 contractions and QR on raw device memory (the ``mps_dev_*`` entry points).

Arrays are passed as torch complex128 tensors on a GPU holding the
column-major (Julia) element order of the reference's arrays, flattened;
``to_dev``/``from_dev`` convert numpy arrays.  torch only provides device
memory here; all arithmetic runs in libmpship.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from mpsamd._lib import call
from mpsamd.core import default_context, _cbuf


def _torch():
    import torch
    return torch


def to_dev(a, device="cuda"):
    torch = _torch()
    flat = np.ascontiguousarray(np.asarray(a, dtype=np.complex128).reshape(-1, order="F"))
    t = torch.from_numpy(flat).to(device)
    torch.cuda.synchronize()   # the library runs on its own stream
    return t


def from_dev(t, shape):
    return t.detach().cpu().numpy().reshape(shape, order="F")


def empty(shape, device="cuda"):
    torch = _torch()
    t = torch.zeros(int(np.prod(shape)), dtype=torch.complex128, device=device)
    torch.cuda.synchronize()
    return t


def _p(t):
    """Device address of a tensor; the caller must keep ``t`` referenced
    until the library call returns (a dropped temporary goes back to torch's
    caching allocator and its memory can be handed to the next operand)."""
    return C.c_void_p(t.data_ptr())


def _ctx(ctx):
    return (ctx or default_context()).h


def prod_AOR(A, O, R, ctx=None):
    """mps_base.jl:7-19 -> [Dl, wl, d, Dr0]."""
    Dl, d, Dr = A.shape
    wl, _, wr, _ = O.shape
    Dr0 = R.shape[2]
    out = empty((Dl, wl, d, Dr0))
    _keep0 = to_dev(A)
    _keep1 = to_dev(O)
    _keep2 = to_dev(R)
    call("mps_dev_prod_AOR", _ctx(ctx), _p(out), _p(_keep0), Dl, d, Dr, _p(_keep1), wl, wr,
         _p(_keep2), Dr0)
    return from_dev(out, (Dl, wl, d, Dr0))


def prod_AOL(A, O, L, ctx=None):
    """mps_base.jl:26-39 -> [Dl0, d, Dr, wr]."""
    Dl, d, Dr = A.shape
    wl, _, wr, _ = O.shape
    Dl0 = L.shape[2]
    out = empty((Dl0, d, Dr, wr))
    _keep0 = to_dev(A)
    _keep1 = to_dev(O)
    _keep2 = to_dev(L)
    call("mps_dev_prod_AOL", _ctx(ctx), _p(out), _p(_keep0), Dl, d, Dr, _p(_keep1), wl, wr,
         _p(_keep2), Dl0)
    return from_dev(out, (Dl0, d, Dr, wr))


def prod_LR(L, R, ctx=None, co=False):
    """mps_base.jl:45-137, dispatched on the ranks of L and R (co: prod_LR_co)."""
    c = _ctx(ctx)
    if L.ndim == 3 and R.ndim == 4:
        r1, r2, r3, r4 = R.shape
        l3 = L.shape[2]
        out = empty((l3, r3, r4))
        _keep0 = to_dev(L)
        _keep1 = to_dev(R)
        call("mps_dev_prod_LR_34", c, _p(out), _p(_keep0), _p(_keep1), r1, r2, l3, r3, r4)
        return from_dev(out, (l3, r3, r4))
    if L.ndim == 4 and R.ndim == 3:
        l1, l2, l3, l4 = L.shape
        r3 = R.shape[2]
        out = empty((l1, l2, r3))
        _keep0 = to_dev(L)
        _keep1 = to_dev(R)
        call("mps_dev_prod_LR_43", c, _p(out), _p(_keep0), _p(_keep1), l1, l2, l3, l4, r3)
        return from_dev(out, (l1, l2, r3))
    if L.ndim == 2 and R.ndim == 3:
        r1, r2, r3 = R.shape
        l1 = L.shape[1] if co else L.shape[0]
        out = empty((l1, r2, r3))
        _keep0 = to_dev(L)
        _keep1 = to_dev(R)
        call("mps_dev_prod_LR_23", c, _p(out), _p(_keep0), _p(_keep1), l1, r1, r2, r3, 1 if co else 0)
        return from_dev(out, (l1, r2, r3))
    if L.ndim == 3 and R.ndim == 2:
        l1, l2, l3 = L.shape
        r2 = R.shape[1]
        out = empty((l1, l2, r2))
        _keep0 = to_dev(L)
        _keep1 = to_dev(R)
        call("mps_dev_prod_LR_32", c, _p(out), _p(_keep0), _p(_keep1), l1, l2, l3, r2)
        return from_dev(out, (l1, l2, r2))
    raise ValueError("prod_LR: unsupported ranks")


def update_re(A0, AOR, ctx=None):
    """mps_base.jl:142-160."""
    r1, r2, r3, r4 = AOR.shape
    a1 = A0.shape[0]
    out = empty((r1, r2, a1))
    _keep0 = to_dev(A0)
    _keep1 = to_dev(AOR)
    call("mps_dev_update_re", _ctx(ctx), _p(out), _p(_keep0), _p(_keep1), r1, r2, r3, r4, a1)
    return from_dev(out, (r1, r2, a1))


def update_le(A0, AOL, ctx=None):
    """mps_base.jl:223-241."""
    l1, l2, l3, l4 = AOL.shape
    a3 = A0.shape[2]
    out = empty((l3, l4, a3))
    _keep0 = to_dev(A0)
    _keep1 = to_dev(AOL)
    call("mps_dev_update_le", _ctx(ctx), _p(out), _p(_keep0), _p(_keep1), l1, l2, l3, l4, a3)
    return from_dev(out, (l3, l4, a3))


def update_renv(A0, AR, ctx=None):
    """mps_base.jl:185-199."""
    a1, d, a2 = A0.shape
    ar1 = AR.shape[0]
    out = empty((ar1, a1))
    _keep0 = to_dev(A0)
    _keep1 = to_dev(AR)
    call("mps_dev_update_renv", _ctx(ctx), _p(out), _p(_keep0), _p(_keep1), a1, d, a2, ar1)
    return from_dev(out, (ar1, a1))


def update_lenv(A0, AL, mode=0, ctx=None):
    """mps_base.jl:267-325 (mode 0: update_lenv, 1: _co, 2: _nc)."""
    a1, d, a3 = A0.shape
    al3 = AL.shape[2]
    shape = (al3, a3) if mode == 1 else (a3, al3)
    out = empty(shape)
    _keep0 = to_dev(A0)
    _keep1 = to_dev(AL)
    call("mps_dev_update_lenv", _ctx(ctx), _p(out), _p(_keep0), _p(_keep1), a1, d, a3, al3, int(mode))
    return from_dev(out, shape)


def local_exp(A0, A, O, ctx=None):
    """mps_base.jl:341-349."""
    a1, d, a3 = A.shape
    res = np.zeros(2)
    keep, p = _cbuf(np.asarray(O, dtype=np.complex128))
    _keep0 = to_dev(A0)
    _keep1 = to_dev(A)
    call("mps_dev_local_exp", _ctx(ctx), res.ctypes.data_as(C.POINTER(C.c_double)), _p(_keep0),
         _p(_keep1), a1, d, a3, p)
    return complex(res[0], res[1])


def qrq(M, ctx=None, transposed_view=False, want_r=False):
    """orthonormalisation.jl:180-198: Q of the Householder QR of an m x n matrix (m >= n).

    With ``transposed_view`` the device array holds M^T column-major and Q is
    written back through the same view (rightorth!'s layout).
    """
    m, n = M.shape
    if transposed_view:
        t = to_dev(np.asarray(M).T)      # element (r,c) at c + n*r
        rs, cs = n, 1
    else:
        t = to_dev(M)
        rs, cs = 1, m
    R = empty((n, n)) if want_r else None
    call("mps_dev_qrq", _ctx(ctx), _p(t), m, n, rs, cs, _p(R) if want_r else None)
    Q = from_dev(t, (n, m)).T if transposed_view else from_dev(t, (m, n))
    if want_r:
        return Q, from_dev(R, (n, n))
    return Q
