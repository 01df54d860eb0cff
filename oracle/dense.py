"""Dense brute-force reference for small chains (TEST INFRASTRUCTURE ONLY).

Independent of the reshape/GEMM chain of ``mps_oracle``: states and operators
are contracted site by site with explicit ``einsum`` index formulas.  The
dense index ordering is column-major over sites (site 1 fastest), matching
``reshape(..., order='F')`` of the reference's column-major arrays.
"""
from __future__ import annotations

import numpy as np


def mps_to_dense(mps):
    """psi[s1, s2, ..., sN] (s1 fastest) as a flat vector."""
    t = mps[0][0]                       # (d, D1)
    t = t.reshape(-1, mps[0].shape[2])
    for A in mps[1:]:
        # t[(prev), a] * A[a, s, b] -> t[(prev, s), b] with prev fastest
        t = np.einsum("pa,asb->psb", t, A).reshape(-1, A.shape[2], order="F")
    return t[:, 0]


def mpo_to_dense(mpo):
    """M[(s_out...), (s_in...)] with site 1 fastest on both sides."""
    W = np.transpose(mpo[0][0], (0, 2, 1))          # (o, i, wr)
    for O in mpo[1:]:
        T = np.einsum("oiw,wsxt->ositx", W, O)      # o, s, i, t, wr
        no, d, ni, d2, wr = T.shape
        W = T.reshape(no * d, ni, d2, wr, order="F").reshape(no * d, ni * d2, wr, order="F")
    return W[:, :, 0]


def site_op(N, d, j, op):
    """Dense operator acting as ``op`` on site j (0-based) of an N-site chain."""
    out = np.ones((1, 1), dtype=np.complex128)
    for k in range(N):
        o = op if k == j else np.eye(d, dtype=np.complex128)
        out = np.kron(o, out)           # site 1 fastest
    return out


def rk4_dense(M, psi, dt):
    """Classic RK4 step of d psi/dt = M psi."""
    k1 = M @ psi
    k2 = M @ (psi + dt / 2 * k1)
    k3 = M @ (psi + dt / 2 * k2)
    k4 = M @ (psi + dt * k3)
    return psi + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
