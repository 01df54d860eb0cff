"""FP64 CPU oracle for the MPS variational-compression / RK4 hot path.

TEST INFRASTRUCTURE ONLY.  This module restates, in numpy + LAPACK, the
algorithm of jdnz/MatrixProductStates.jl that the HIP library replaces.  It is
imported by ``tests/``, by ``__graft_entry__.smoke()`` (as the checker) and by
``bench.py``'s ``cpu_baseline`` leg (as the timed CPU baseline).  The product
path (``mpsamd``) never imports it.

Conventions (identical to the reference, which is column-major Julia):

* every reshape is column-major (``order='F'``), every ``permutedims`` keeps
  the reference's 1-based permutation vector;
* MPS site ``A[Dl, sigma, Dr]``; MPO site ``O[wl, sigma_out, wr, sigma_in]``;
* operator environment ``E[Dket, w, Dbra]``; LEFT operator environments are
  stored conjugated (``mps_base.jl:219-241``);
* overlap environments: left ``L[Dbra, Dket]``, right ``R[Dket, Dbra]``
  (``mps_base.jl:185-218, 264-281``);
* lists of sites / environment slots are 0-based here: Julia ``env[jj]`` is
  ``env[jj - 1]``, so site ``j`` has its left environment in ``env[j]`` and its
  right environment in ``env[j + 1]``.

QR follows the reference exactly: ``qrq!`` is LAPACK ``zgeqrf`` followed by
``zungqr`` (``orthonormalisation.jl:180-184``), called here through
``scipy.linalg.lapack``.  Julia's ``qr`` (used by ``leftorth``/``rightorth``)
computes the same Householder factorisation through ``zgeqrt``; the oracle
uses ``zgeqrf`` for both, which differs from ``zgeqrt`` only in rounding.

Pinning status of every function is listed in the test files
(``tests/test_oracle_*.py``): contractions against brute-force ``einsum`` of
their index formulas, sweeps against dense linear algebra at full bond
dimension, RK4 against dense RK4, QR against LAPACK invariants and the
closed form for zero-padded inputs.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lapack, expm

CDT = np.complex128


def _r(x, *shape):
    """Julia ``reshape`` (column-major)."""
    return np.reshape(x, shape, order="F")


def _perm(x, perm):
    """Julia ``permutedims`` with a 1-based permutation vector."""
    return np.transpose(x, [p - 1 for p in perm])


def _H(x):
    return x.conj().T


# --------------------------------------------------------------------------
# L1 contractions -- src/mps_base.jl
# --------------------------------------------------------------------------

def prod_AOR(A, O, R):
    """mps_base.jl:7-19.  AOR[a,w,s',c] = sum A[a,s,a3] R[a3,w',c] O[w,s',w',s]."""
    a1, a2, a3 = A.shape
    _, r2, r3 = R.shape
    o1 = O.shape[0]
    AOR = _r(A, a1 * a2, a3) @ _r(R, a3, r2 * r3)
    AOR = _perm(_r(AOR, a1, a2, r2, r3), [3, 2, 1, 4])
    AOR = _r(O, o1 * a2, r2 * a2) @ _r(AOR, r2 * a2, a1 * r3)
    return _perm(_r(AOR, o1, a2, a1, r3), [3, 1, 2, 4])


def prod_AOL(A, O, L):
    """mps_base.jl:26-39.  AOL[b,s',a',w'] = sum conj(L[a,w,b]) A[a,s,a'] O[w,s',w',s]."""
    a1, d, a3 = A.shape
    _, l2, l3 = L.shape
    o3 = O.shape[2]
    AOL = _H(_r(L, a1, l2 * l3)) @ _r(A, a1, d * a3)
    AOL = _perm(_r(AOL, l2, l3, d, a3), [2, 4, 1, 3])
    AOL = _r(AOL, l3 * a3, l2 * d)
    AOL = AOL @ _r(_perm(O, [1, 4, 2, 3]), l2 * d, d * o3)
    return _perm(_r(AOL, l3, a3, d, o3), [1, 3, 2, 4])


def prod_LR_34(L, R):
    """mps_base.jl:45-61 (3-index L, 4-index R): L^H * R -> [l3, r3, r4]."""
    r1, r2, r3, r4 = R.shape
    l3 = L.shape[2]
    return _r(_H(_r(L, r1 * r2, l3)) @ _r(R, r1 * r2, r3 * r4), l3, r3, r4)


def prod_LR_43(L, R):
    """mps_base.jl:65-81 (4-index L, 3-index R): L * R -> [l1, l2, r3]."""
    l1, l2, l3, l4 = L.shape
    r3 = R.shape[2]
    return _r(_r(L, l1 * l2, l3 * l4) @ _r(R, l3 * l4, r3), l1, l2, r3)


def prod_LR_23(L, R):
    """mps_base.jl:85-90, 100-108 (2-index L, 3-index R): L * R."""
    r1, r2, r3 = R.shape
    return _r(L @ _r(R, r1, r2 * r3), L.shape[0], r2, r3)


def prod_LR_co_23(L, R):
    """mps_base.jl:92-97, 110-118: L^T * R (no conjugation)."""
    r1, r2, r3 = R.shape
    return _r(L.T @ _r(R, r1, r2 * r3), L.shape[1], r2, r3)


def prod_LR_32(L, R):
    """mps_base.jl:122-137 (3-index L, 2-index R): L * R."""
    l1, l2, l3 = L.shape
    return _r(_r(L, l1 * l2, l3) @ R, l1, l2, R.shape[1])


def prod_LR(L, R):
    """Julia multiple dispatch of ``prod_LR`` on the ranks of its arguments."""
    key = (L.ndim, R.ndim)
    return {(3, 4): prod_LR_34, (4, 3): prod_LR_43, (2, 3): prod_LR_23,
            (3, 2): prod_LR_32}[key](L, R)


def update_re(A0, AOR):
    """mps_base.jl:142-160.  R'[a,w,c_l] = sum AOR[a,w,s',c] conj(A0[c_l,s',c])."""
    r1, r2, r3, r4 = AOR.shape
    a1 = A0.shape[0]
    return _r(_r(AOR, r1 * r2, r3 * r4) @ _H(_r(A0, a1, r3 * r4)), r1, r2, a1)


def update_re_full(A0, O, A, FR):
    """mps_base.jl:164-180."""
    return update_re(A0, prod_AOR(A, O, FR))


def update_renv(A0, AR):
    """mps_base.jl:185-199.  R'[a,c_l] = sum AR[a,s,c] conj(A0[c_l,s,c])."""
    a1, d, a2 = A0.shape
    return _r(AR, AR.shape[0], d * a2) @ _H(_r(A0, a1, d * a2))


def update_renv_full(A0, A, FR):
    """mps_base.jl:204-218."""
    return update_renv(A0, prod_LR_32(A, FR))


def update_le(A0, AOL):
    """mps_base.jl:223-241 (conjugated left env).  L'[a',w',c] = sum conj(AOL[b,s',a',w']) A0[b,s',c]."""
    l1, l2, l3, l4 = AOL.shape
    a3 = A0.shape[2]
    return _r(_H(_r(AOL, l1 * l2, l3 * l4)) @ _r(A0, l1 * l2, a3), l3, l4, a3)


def update_le_full(A0, O, A, FL):
    """mps_base.jl:246-262."""
    return update_le(A0, prod_AOL(A, O, FL))


def update_lenv(A0, AL):
    """mps_base.jl:267-281.  L'[c,a'] = sum conj(A0[b,s,c]) AL[b,s,a']."""
    a1, d, a3 = A0.shape
    return _H(_r(A0, a1 * d, a3)) @ _r(AL, a1 * d, AL.shape[2])


def update_lenv_full(A0, A, L):
    """mps_base.jl:293-307."""
    return update_lenv(A0, prod_LR_23(L, A))


def update_lenv_co(A0, AL):
    """mps_base.jl:283-289.  L'[a',c] = sum AL[b,s,a'] conj(A0[b,s,c])  (L[ket,bra])."""
    a1, d, a3 = A0.shape
    return _r(AL, a1 * d, AL.shape[2]).T @ _r(A0.conj(), a1 * d, a3)


def update_lenv_co_full(A0, A, L):
    """mps_base.jl:309-315."""
    return update_lenv_co(A0, prod_LR_co_23(L, A))


def update_lenv_nc(A0, AL):
    """mps_base.jl:319-325 (no conjugation of A0)."""
    a1, d, a3 = A0.shape
    return _r(A0, a1 * d, a3).T @ _r(AL, a1 * d, AL.shape[2])


def update_lenv_nc_full(A0, A, L):
    """mps_base.jl:330-336."""
    return update_lenv_nc(A0, prod_LR_23(L, A))


def local_exp(A0, A, O):
    """mps_base.jl:341-349.  trace(O * (A_p A0_p^H))."""
    a1, d, a3 = A.shape
    Ap = _r(_perm(A, [2, 1, 3]), d, a1 * a3)
    A0p = _r(_perm(A0, [2, 1, 3]), d, a1 * a3)
    return np.trace(O @ (Ap @ _H(A0p)))


# --------------------------------------------------------------------------
# L2 gauge moves -- src/orthonormalisation.jl
# --------------------------------------------------------------------------

# Rank-revealing cutoff of the Householder generator (0 = exact LAPACK
# semantics).  With a cutoff tol > 0 a reflector whose trailing column norm is
# <= tol * ||A||_F is taken as the identity (tau = 0), the same branch zlarfg
# takes for an exactly zero column.  The device QR implements the same rule.
QR_RANK_TOL = 0.0


def set_qr_rank_tol(tol):
    global QR_RANK_TOL
    QR_RANK_TOL = float(tol)


def zlarfg(alpha, x, thr=0.0):
    """LAPACK zlarfg (Householder generator) restated; returns (tau, beta, v_tail).

    H = I - tau v v^H with v = [1, v_tail] and H^H [alpha; x] = [beta; 0],
    beta real, beta = -sign(Re alpha) * ||[alpha; x]||.  tau = 0 when x = 0
    and Im alpha = 0 (or, with thr > 0, both below thr).
    """
    xnorm = float(np.linalg.norm(x))
    if thr > 0.0:
        if xnorm <= thr and abs(alpha.imag) <= thr:
            return 0j, alpha, x
    elif xnorm == 0.0 and alpha.imag == 0.0:
        return 0j, alpha, x
    beta = -np.copysign(np.sqrt(alpha.real ** 2 + alpha.imag ** 2 + xnorm ** 2), alpha.real)
    tau = complex((beta - alpha.real) / beta, -alpha.imag / beta)
    return tau, complex(beta), x * (1.0 / (alpha - beta))


def householder_geqr2(A, tol=0.0):
    """LAPACK zgeqr2 restated (unblocked): packed QR and tau."""
    A = np.array(A, dtype=CDT, order="F")
    m, n = A.shape
    thr = tol * float(np.linalg.norm(A)) if tol > 0 else 0.0
    k = min(m, n)
    tau = np.zeros(k, dtype=CDT)
    for i in range(k):
        t, beta, xv = zlarfg(complex(A[i, i]), A[i + 1:, i], thr)
        tau[i] = t
        A[i + 1:, i] = xv
        if i < n - 1 and t != 0:
            v = np.concatenate([[1.0 + 0j], A[i + 1:, i]])
            w = v.conj() @ A[i:, i + 1:]
            A[i:, i + 1:] -= np.conj(t) * np.outer(v, w)
        A[i, i] = beta
    return A, tau


def householder_ung2r(A, tau):
    """LAPACK zung2r restated: the m x n Q = H(1)...H(k) I[:, :n]."""
    A = np.array(A, dtype=CDT, order="F")
    m, n = A.shape
    k = len(tau)
    for j in range(k, n):
        A[:, j] = 0
        A[j, j] = 1
    for i in range(k - 1, -1, -1):
        if i < n - 1:
            A[i, i] = 1.0
            if tau[i] != 0:
                v = A[i:, i].copy()
                w = v.conj() @ A[i:, i + 1:]
                A[i:, i + 1:] -= tau[i] * np.outer(v, w)
        if i < m - 1:
            A[i + 1:, i] *= -tau[i]
        A[i, i] = 1.0 - tau[i]
        A[:i, i] = 0
    return A


def geqrf(M):
    """LAPACK zgeqrf: returns (packed QR, tau)."""
    if QR_RANK_TOL > 0.0:
        return householder_geqr2(M, QR_RANK_TOL)
    qr, tau, _, info = lapack.zgeqrf(np.asfortranarray(M, dtype=CDT))
    if info != 0:
        raise RuntimeError(f"zgeqrf info={info}")
    return qr, tau


def _ungqr(qr, tau):
    if QR_RANK_TOL > 0.0:
        return householder_ung2r(qr, tau)
    q, _, info = lapack.zungqr(np.asfortranarray(qr), tau)
    if info != 0:
        raise RuntimeError(f"zungqr info={info}")
    return q


def qrq(M):
    """orthonormalisation.jl:180-184: ``LAPACK.orgqr!(LAPACK.geqrf!(A)...)``.

    Returns the m x n matrix Q (R is discarded).  Requires m >= n, which the
    bond-dimension rule (constructors.jl:15-18) guarantees on the hot path.
    """
    m, n = M.shape
    if m < n:
        raise ValueError("qrq: m < n is not produced by the reference path")
    qr, tau = geqrf(M)
    return _ungqr(qr, tau)


def thin_qr(M):
    """Julia 0.6 ``qr(A)`` (thin Q and R, Householder with LAPACK conventions)."""
    m, n = M.shape
    k = min(m, n)
    qr, tau = geqrf(M)
    R = np.triu(qr[:k, :])
    return _ungqr(qr[:, :k], tau[:k]), R


def leftorth_(A):
    """orthonormalisation.jl:164-171 (``leftorth!``): Q of reshape(A, d*Dl, Dr)."""
    Dl, d, Dr = A.shape
    Q = qrq(_r(A, d * Dl, Dr))
    return _r(Q, Dl, d, Q.shape[1])


def rightorth_(A):
    """orthonormalisation.jl:227-234 (``rightorth!``): QR of the plain transpose."""
    Dl, d, Dr = A.shape
    Q = qrq(_r(A, Dl, d * Dr).T)
    return _r(Q.T, Q.shape[1], d, Dr)


def leftorth(A):
    """orthonormalisation.jl:150-157: (AL, C) with A = AL * C."""
    Dl, d, Dr = A.shape
    Q, C = thin_qr(_r(A, d * Dl, Dr))
    return _r(Q, Dl, d, Q.shape[1]), C


def rightorth(A):
    """orthonormalisation.jl:204-211: (AR, C) with A = C * AR."""
    Dl, d, Dr = A.shape
    Q, C = thin_qr(_r(A, Dl, d * Dr).T)
    return _r(Q.T, Q.shape[1], d, Dr), C.T


def leftorthogonalizeQR_(mps):
    """orthonormalisation.jl:43-56.  Returns the norm^2 (not normalised)."""
    n = len(mps)
    for i in range(n - 1):
        mps[i], r = leftorth(mps[i])
        mps[i + 1] = prod_LR_23(r, mps[i + 1])
    return float(np.sum(np.abs(mps[n - 1]) ** 2))


def leftorthonormalizationQR_(mps):
    """orthonormalisation.jl:29-36."""
    no = leftorthogonalizeQR_(mps)
    mps[-1] = mps[-1] / np.sqrt(no)
    return no


def rightorthogonalizeQR_(mps):
    """orthonormalisation.jl:77-90."""
    n = len(mps)
    for i in range(n - 1, 0, -1):
        mps[i], r = rightorth(mps[i])
        mps[i - 1] = prod_LR_32(mps[i - 1], r)
    return float(np.sum(np.abs(mps[0]) ** 2))


def rightorthonormalizationQR_(mps):
    """orthonormalisation.jl:64-71."""
    no = rightorthogonalizeQR_(mps)
    mps[0] = mps[0] / np.sqrt(no)
    return no


def normalize_lo_(mps):
    """orthonormalisation.jl:97-103: norm^2 from the last site, then rescale it."""
    no = float(np.sum(np.abs(mps[-1]) ** 2))
    mps[-1] = mps[-1] / np.sqrt(no)
    return no


def norm_lo(mps):
    """orthonormalisation.jl:117-121."""
    return float(np.sum(np.abs(mps[-1]) ** 2))


# --------------------------------------------------------------------------
# constructors -- src/constructors.jl
# --------------------------------------------------------------------------

def _dims_update(dims, n, na, d):
    # constructors.jl:15-18 (dims is 1-based: dims[1..na+1], dims[0] unused)
    dims[n + 1] = min(dims[n] * d, dims[n + 1])
    dims[na + 1 - n] = min(dims[na + 2 - n] * d, dims[na + 1 - n])


def mpsproductstate(na, dmax, d, state):
    """constructors.jl:5-61.  ``state`` is a length-d vector or ``n -> vector`` (1-based n)."""
    dims = np.ones(na + 2, dtype=np.int64)
    dims[2:na + 1] = dmax
    r = np.ones((1, 1), dtype=CDT)
    mps = [None] * na
    st = state if callable(state) else (lambda n: state)
    for n in range(1, na):
        _dims_update(dims, n, na, d)
        site = np.zeros((dims[n], d, dims[n + 1]), dtype=CDT)
        site[0, :, 0] = st(n)
        site = prod_LR_23(r, site)
        mps[n - 1], r = leftorth(site)
    site = np.zeros((dims[na], d, dims[na + 1]), dtype=CDT)
    site[0, :, 0] = st(na)
    mps[na - 1] = prod_LR_23(r, site)
    return mps, [int(x) for x in dims[1:]]


def mpsgroundstate(na, dmax, d):
    """constructors.jl:67-75: product state with local state e_1."""
    state = np.zeros(d, dtype=CDT)
    state[0] = 1.0
    return mpsproductstate(na, dmax, d, state)


def crandn(rng, shape):
    """MatrixProductStates.jl:26-27: complex normal with the sqrt(1/2) scaling."""
    s = np.sqrt(0.5)
    return (s * rng.standard_normal(shape) + 1j * s * rng.standard_normal(shape)).astype(CDT)


def mpsrandom(na, dmax, d, rng):
    """constructors.jl:111-129: random left-orthonormal MPS (numpy RNG, seeded by the caller)."""
    dims = np.ones(na + 2, dtype=np.int64)
    dims[2:na + 1] = dmax
    mps = [None] * na
    for n in range(1, na + 1):
        dims[n + 1] = min(dims[n] * d, dims[n + 1])
        dims[na + 1 - n] = min(dims[na + 2 - n] * d, dims[na + 1 - n])
        mps[n - 1] = leftorth_(crandn(rng, (dims[n], d, dims[n + 1])))
    return mps, [int(x) for x in dims[1:]]


def makempo(mpomat, na, d):
    """constructors.jl:229-253.  ``mpomat[l][m]`` is ``n -> d x d`` (1-based n)."""
    b1, b2 = len(mpomat), len(mpomat[0])
    mpo = []
    for n in range(1, na + 1):
        site = np.zeros((b1, d, b2, d), dtype=CDT)
        for m in range(b2):
            for l in range(b1):
                site[l, :, m, :] = mpomat[l][m](n)
        mpo.append(site)
    mpo[0] = mpo[0][0:1].copy()
    mpo[-1] = mpo[-1][:, :, b2 - 1:b2, :].copy()
    return mpo


def makemps(mpsmat, na, d):
    """constructors.jl:198-222.  ``mpsmat[l][m]`` is ``n -> length-d vector``."""
    b1, b2 = len(mpsmat), len(mpsmat[0])
    mps = []
    for n in range(1, na + 1):
        site = np.zeros((b1, d, b2), dtype=CDT)
        for m in range(b2):
            for l in range(b1):
                site[l, :, m] = mpsmat[l][m](n)
        mps.append(site)
    mps[0] = mps[0][0:1].copy()
    mps[-1] = mps[-1][:, :, b2 - 1:b2].copy()
    return mps


def build_env_op(dims1, dims2, mpo_dims):
    """constructors.jl:347-359 (3-index operator environments, N+1 slots)."""
    n = len(dims1)
    env = [np.ones((1, 1, 1), dtype=CDT)]
    for k in range(1, n - 1):
        env.append(np.zeros((dims1[k], int(mpo_dims[k]), dims2[k]), dtype=CDT))
    env.append(np.ones((1, 1, 1), dtype=CDT))
    return env


def build_env(dims1, dims2):
    """constructors.jl:363-375 (2-index overlap environments, N+1 slots)."""
    n = len(dims1)
    env = [np.ones((1, 1), dtype=CDT)]
    for k in range(1, n - 1):
        env.append(np.zeros((dims1[k], dims2[k]), dtype=CDT))
    env.append(np.ones((1, 1), dtype=CDT))
    return env


def mpo_to_mps_site(mpo):
    """constructors.jl:310-315."""
    dl, d, dr, _ = mpo.shape
    return _r(_perm(mpo, [1, 2, 4, 3]), dl, d * d, dr)


def mpo_to_mps(mpo):
    """constructors.jl:298-308."""
    return [mpo_to_mps_site(m) for m in mpo]


# --------------------------------------------------------------------------
# MPO / MPS products and scalar products -- src/MatrixProductStates.jl
# --------------------------------------------------------------------------

def apply_site_MPOtoMPO(mpo1, mpo2):
    """MatrixProductStates.jl:89-98: site of the operator product mpo2*mpo1."""
    dwl1, d, dwr1, _ = mpo1.shape
    dwl2, _, dwr2, _ = mpo2.shape
    ai = _r(mpo2, dwl2 * d * dwr2, d) @ _r(_perm(mpo1, [2, 1, 3, 4]), d, dwl1 * dwr1 * d)
    ai = _perm(_r(ai, dwl2, d, dwr2, dwl1, dwr1, d), [4, 1, 2, 5, 3, 6])
    return _r(ai, dwl1 * dwl2, d, dwr1 * dwr2, d)


def applyMPOtoMPO(mpo1, mpo2):
    """MatrixProductStates.jl:67-76."""
    return [apply_site_MPOtoMPO(a, b) for a, b in zip(mpo1, mpo2)]


def apply_site_MPOtoMPS(ms, mo):
    """MatrixProductStates.jl:126-133: ai[a,wl,s,b,wr] = sum ms[a,t,b] mo[wl,s,wr,t]."""
    dl, _, dr = ms.shape
    dwl, d, dwr, _ = mo.shape
    ai = np.einsum("atb,lsrt->alsbr", ms, mo)
    return _r(ai, dl * dwl, d, dr * dwr)


def applyMPOtoMPS(mps, mpo):
    """MatrixProductStates.jl:115-124."""
    return [apply_site_MPOtoMPS(a, o) for a, o in zip(mps, mpo)]


def apply_site_operator_(site, op):
    """MatrixProductStates.jl:135-143: site[:,s,:] <- sum_t op[s,t] site[:,t,:] (in place)."""
    m1, m2, m3 = site.shape
    t = _r(_perm(site, [2, 1, 3]), m2, m1 * m3)
    t = op @ t
    site[...] = _perm(_r(t, m2, m1, m3), [2, 1, 3])
    return site


def conj_site_mpo(m):
    """MatrixProductStates.jl:409-413."""
    return _perm(m.conj(), [1, 4, 3, 2])


def conj_mpo(mpo):
    """MatrixProductStates.jl:395-405."""
    return [conj_site_mpo(m) for m in mpo]


def scal_op_prod(mps1, op, mps2):
    """MatrixProductStates.jl:327-340.  Returns conj(<mps1|op|mps2>) (conjugated left env)."""
    lenv = np.ones((1, 1, 1), dtype=CDT)
    for a1, o, a2 in zip(mps1, op, mps2):
        lenv = update_le_full(a1, o, a2, lenv)
    return complex(lenv.flat[0])


def scal_prod(mps1, mps2):
    """MatrixProductStates.jl:345-358: <mps1|mps2>."""
    lenv = np.ones((1, 1), dtype=CDT)
    for a1, a2 in zip(mps1, mps2):
        lenv = update_lenv_full(a1, a2, lenv)
    return complex(lenv.flat[0])


def scal_prod_no_conj(mps1, mps2):
    """MatrixProductStates.jl:366-379: bilinear sum mps1 . mps2."""
    lenv = np.ones((1, 1), dtype=CDT)
    for a1, a2 in zip(mps1, mps2):
        lenv = update_lenv_nc_full(a1, a2, lenv)
    return complex(lenv.flat[0])


def normMPS(mps):
    """MatrixProductStates.jl:317-322."""
    return abs(scal_prod(mps, mps))


def diff_norm(mps1, mps2):
    """MatrixProductStates.jl:388-392."""
    return normMPS(mps1) + normMPS(mps2) - 2 * scal_prod(mps1, mps2).real


def mpsdims(mps):
    """MatrixProductStates.jl:41-51."""
    dims = [int(a.shape[0]) for a in mps]
    return dims + [1]


# --------------------------------------------------------------------------
# L3 sweeps -- src/compression.jl
# --------------------------------------------------------------------------

def compress_var_(A0, A, env, maxsweeps):
    """compression.jl:65-103: overlap-only one-site fit of A into A0 (A0 left canonical)."""
    na = len(A0)
    env[0] = np.ones((1, 1), dtype=CDT)
    for j in range(na - 1):
        env[j + 1] = update_lenv_co_full(A0[j], A[j], env[j])
    env[na] = np.ones((1, 1), dtype=CDT)
    for _ in range(maxsweeps):
        for j in range(na - 1, 0, -1):
            AR = prod_LR_32(A[j], env[j + 1])
            A0[j] = prod_LR_co_23(env[j], AR)
            A0[j] = rightorth_(A0[j])
            env[j] = update_renv(A0[j], AR)
        AR = prod_LR_32(A[0], env[1])
        A0[0] = prod_LR_23(env[0], AR)
        for j in range(na - 1):
            AL = prod_LR_co_23(env[j], A[j])
            A0[j] = prod_LR_32(AL, env[j + 1])
            A0[j] = leftorth_(A0[j])
            env[j + 1] = update_lenv_co(A0[j], AL)
        AL = prod_LR_co_23(env[na - 1], A[na - 1])
        A0[na - 1] = prod_LR_32(AL, env[na])


def compress_var_apply_H_(A0, envop, Ain, A, H, maxsweeps):
    """compression.jl:217-258: ``maxsweeps`` full sweeps of A0 ~ H*A (envs built from Ain)."""
    na = len(Ain)
    envop[0] = np.ones((1, 1, 1), dtype=CDT)
    for j in range(na - 1):
        envop[j + 1] = update_le_full(Ain[j], H[j], A[j], envop[j])
    envop[na] = np.ones((1, 1, 1), dtype=CDT)
    for _ in range(maxsweeps):
        for j in range(na - 1, 0, -1):
            AOR = prod_AOR(A[j], H[j], envop[j + 1])
            A0[j] = prod_LR_34(envop[j], AOR)
            A0[j] = rightorth_(A0[j])
            envop[j] = update_re(A0[j], AOR)
        for j in range(na - 1):
            AOL = prod_AOL(A[j], H[j], envop[j])
            A0[j] = prod_LR_43(AOL, envop[j + 1])
            A0[j] = leftorth_(A0[j])
            envop[j + 1] = update_le(A0[j], AOL)
        AOL = prod_AOL(A[na - 1], H[na - 1], envop[na - 1])
        A0[na - 1] = prod_LR_43(AOL, envop[na])


def compress_var_apply_H_left_(A0, envop, Ain, A, H):
    """compression.jl:321-347: RK4 stage 1, one right-to-left half sweep of A0 ~ H*A."""
    na = len(Ain)
    envop[0] = np.ones((1, 1, 1), dtype=CDT)
    for j in range(na - 1):
        envop[j + 1] = update_le_full(Ain[j], H[j], A[j], envop[j])
    envop[na] = np.ones((1, 1, 1), dtype=CDT)
    for j in range(na - 1, 0, -1):
        AOR = prod_AOR(A[j], H[j], envop[j + 1])
        A0[j] = prod_LR_34(envop[j], AOR)
        A0[j] = rightorth_(A0[j])
        envop[j] = update_re(A0[j], AOR)
    AOR = prod_AOR(A[0], H[0], envop[1])
    A0[0] = prod_LR_34(envop[0], AOR)


def compress_sum_var_apply_H_left_(A0, env, envop, Ain, coef, A, H):
    """compression.jl:889-944: right-to-left half sweep of sum_l c_l A_l + c_numv H A_numv."""
    na = len(Ain)
    numv = len(coef)
    for ll in range(numv - 1):
        env[ll][0] = np.ones((1, 1), dtype=CDT)
        for j in range(na - 1):
            env[ll][j + 1] = update_lenv_full(Ain[j], A[ll][j], env[ll][j])
        env[ll][na] = np.full((1, 1), coef[ll], dtype=CDT)
    envop[0] = np.ones((1, 1, 1), dtype=CDT)
    for j in range(na - 1):
        envop[j + 1] = update_le_full(Ain[j], H[j], A[numv - 1][j], envop[j])
    envop[na] = np.full((1, 1, 1), coef[numv - 1], dtype=CDT)
    AR = [None] * (numv - 1)
    for j in range(na - 1, 0, -1):
        AR[0] = prod_LR_32(A[0][j], env[0][j + 1])
        t = prod_LR_23(env[0][j], AR[0])
        for ll in range(1, numv - 1):
            AR[ll] = prod_LR_32(A[ll][j], env[ll][j + 1])
            t = t + prod_LR_23(env[ll][j], AR[ll])
        AOR = prod_AOR(A[numv - 1][j], H[j], envop[j + 1])
        t = t + prod_LR_34(envop[j], AOR)
        A0[j] = rightorth_(t)
        for ll in range(numv - 1):
            env[ll][j] = update_renv(A0[j], AR[ll])
        envop[j] = update_re(A0[j], AOR)
    AR[0] = prod_LR_32(A[0][0], env[0][1])
    t = prod_LR_23(env[0][0], AR[0])
    for ll in range(1, numv - 1):
        AR[ll] = prod_LR_32(A[ll][0], env[ll][1])
        t = t + prod_LR_23(env[ll][0], AR[ll])
    AOR = prod_AOR(A[numv - 1][0], H[0], envop[1])
    A0[0] = t + prod_LR_34(envop[0], AOR)


def compress_sum_var_apply_H_right_(A0, env, envop, Ain, coef, A, H):
    """compression.jl:1007-1063: left-to-right half sweep (RK4 stages 2 and 4).

    ``A0`` may be the same list as ``A[0]`` (the 10-argument RK4 of
    timeevolution.jl:178-186): site j of ``A[0]`` is read before it is
    replaced, exactly as in the reference.
    """
    na = len(Ain)
    numv = len(coef)
    for ll in range(numv - 1):
        env[ll][na] = np.ones((1, 1), dtype=CDT)
        for j in range(na - 1, 0, -1):
            env[ll][j] = update_renv_full(Ain[j], A[ll][j], env[ll][j + 1])
        env[ll][0] = np.full((1, 1), coef[ll], dtype=CDT)
    envop[na] = np.ones((1, 1, 1), dtype=CDT)
    for j in range(na - 1, 0, -1):
        envop[j] = update_re_full(Ain[j], H[j], A[numv - 1][j], envop[j + 1])
    envop[0] = np.full((1, 1, 1), coef[numv - 1], dtype=CDT)
    AL = [None] * (numv - 1)
    for j in range(na - 1):
        AL[0] = prod_LR_23(env[0][j], A[0][j])
        t = prod_LR_32(AL[0], env[0][j + 1])
        for ll in range(1, numv - 1):
            AL[ll] = prod_LR_23(env[ll][j], A[ll][j])
            t = t + prod_LR_32(AL[ll], env[ll][j + 1])
        AOL = prod_AOL(A[numv - 1][j], H[j], envop[j])
        t = t + prod_LR_43(AOL, envop[j + 1])
        A0[j] = leftorth_(t)
        for ll in range(numv - 1):
            env[ll][j + 1] = update_lenv(A0[j], AL[ll])
        envop[j + 1] = update_le(A0[j], AOL)
    j = na - 1
    AL[0] = prod_LR_23(env[0][j], A[0][j])
    t = prod_LR_32(AL[0], env[0][j + 1])
    for ll in range(1, numv - 1):
        AL[ll] = prod_LR_23(env[ll][j], A[ll][j])
        t = t + prod_LR_32(AL[ll], env[ll][j + 1])
    AOL = prod_AOL(A[numv - 1][j], H[j], envop[j])
    A0[j] = t + prod_LR_43(AOL, envop[j + 1])


def compress_sum_var(Ain, coef, A, maxsweeps):
    """compression.jl:496-565: allocating overlap-only fit of sum_l coef_l A_l."""
    na = len(Ain)
    numv = len(coef)
    FL = [[None] * na for _ in range(numv)]
    FR = [[None] * na for _ in range(numv)]
    A0 = [None] * na
    AR = [None] * numv
    for ll in range(numv):
        FL[ll][0] = np.ones((1, 1), dtype=CDT)
        for j in range(na - 1):
            FL[ll][j + 1] = update_lenv_full(Ain[j], A[ll][j], FL[ll][j])
        FR[ll][na - 1] = coef[ll] * np.ones((1, 1), dtype=CDT)
    for _ in range(maxsweeps):
        for j in range(na - 1, 0, -1):
            AR[0] = prod_LR_32(A[0][j], FR[0][j])
            t = prod_LR_23(FL[0][j], AR[0])
            for ll in range(1, numv):
                AR[ll] = prod_LR_32(A[ll][j], FR[ll][j])
                t = t + prod_LR_23(FL[ll][j], AR[ll])
            A0[j] = rightorth_(t)
            for ll in range(numv):
                FR[ll][j - 1] = update_renv(A0[j], AR[ll])
        for j in range(na - 1):
            AR[0] = prod_LR_23(FL[0][j], A[0][j])
            t = prod_LR_32(AR[0], FR[0][j])
            for ll in range(1, numv):
                AR[ll] = prod_LR_23(FL[ll][j], A[ll][j])
                t = t + prod_LR_32(AR[ll], FR[ll][j])
            A0[j] = leftorth_(t)
            for ll in range(numv):
                FL[ll][j + 1] = update_lenv(A0[j], AR[ll])
    j = na - 1
    AR[0] = prod_LR_23(FL[0][j], A[0][j])
    t = prod_LR_32(AR[0], FR[0][j])
    for ll in range(1, numv):
        AR[ll] = prod_LR_23(FL[ll][j], A[ll][j])
        t = t + prod_LR_32(AR[ll], FR[ll][j])
    A0[j] = t
    return A0


def compress_sum_var_apply_H_(A0, envop, coef, A, H, maxsweeps):
    """compression.jl:805-887, restated by intent.

    As written the reference writes ``env[ll][1][1]`` into an unassigned
    array slot (compression.jl:816-817) and throws.  The intent -- overlap
    environments starting from 1 on the left and ``coef[ll]`` on the right,
    exactly as in ``compress_sum_var_apply_H_left!`` -- is what is restated.
    """
    na = len(A0)
    numv = len(coef)
    env = [[None] * (na + 1) for _ in range(numv - 1)]
    for ll in range(numv - 1):
        env[ll][0] = np.ones((1, 1), dtype=CDT)
        for j in range(na - 1):
            env[ll][j + 1] = update_lenv_full(A0[j], A[ll][j], env[ll][j])
        env[ll][na] = np.full((1, 1), coef[ll], dtype=CDT)
    envop[0] = np.ones((1, 1, 1), dtype=CDT)
    for j in range(na - 1):
        envop[j + 1] = update_le_full(A0[j], H[j], A[numv - 1][j], envop[j])
    envop[na] = np.full((1, 1, 1), coef[numv - 1], dtype=CDT)
    AR = [None] * (numv - 1)
    for _ in range(maxsweeps):
        for j in range(na - 1, 0, -1):
            AR[0] = prod_LR_32(A[0][j], env[0][j + 1])
            t = prod_LR_23(env[0][j], AR[0])
            for ll in range(1, numv - 1):
                AR[ll] = prod_LR_32(A[ll][j], env[ll][j + 1])
                t = t + prod_LR_23(env[ll][j], AR[ll])
            AOR = prod_AOR(A[numv - 1][j], H[j], envop[j + 1])
            t = t + prod_LR_34(envop[j], AOR)
            A0[j] = rightorth_(t)
            for ll in range(numv - 1):
                env[ll][j] = update_renv(A0[j], AR[ll])
            envop[j] = update_re(A0[j], AOR)
        for j in range(na - 1):
            AR[0] = prod_LR_23(env[0][j], A[0][j])
            t = prod_LR_32(AR[0], env[0][j + 1])
            for ll in range(1, numv - 1):
                AR[ll] = prod_LR_23(env[ll][j], A[ll][j])
                t = t + prod_LR_32(AR[ll], env[ll][j + 1])
            AOL = prod_AOL(A[numv - 1][j], H[j], envop[j])
            t = t + prod_LR_43(AOL, envop[j + 1])
            A0[j] = leftorth_(t)
            for ll in range(numv - 1):
                env[ll][j + 1] = update_lenv(A0[j], AR[ll])
            envop[j + 1] = update_le(A0[j], AOL)
    j = na - 1
    AR[0] = prod_LR_23(env[0][j], A[0][j])
    t = prod_LR_32(AR[0], env[0][j + 1])
    for ll in range(1, numv - 1):
        AR[ll] = prod_LR_23(env[ll][j], A[ll][j])
        t = t + prod_LR_32(AR[ll], env[ll][j + 1])
    AOL = prod_AOL(A[numv - 1][j], H[j], envop[j])
    A0[j] = t + prod_LR_43(AOL, envop[j + 1])


def compressionSVD_(mps, dmax, tol=1e-15):
    """compression.jl:1280-1312: right-to-left SVD truncation of a left-canonical MPS.

    Returns ``(err_glob, err_trim, ent_ent)`` and leaves ``mps`` right canonical.
    """
    n = len(mps)
    err_trim = np.zeros(n)
    ent_ent = np.zeros(n)
    dnew = 1
    r = np.ones((1, 1), dtype=CDT)
    for i in range(n - 1, -1, -1):
        dl, d, dr = mps[i].shape
        site = _r(mps[i], dl * d, dr) @ r
        u, s, vh = np.linalg.svd(_r(site, dl, dnew * d), full_matrices=False)
        keep = s > 0
        s = s[keep]
        u = u[:, keep]
        vh = vh[keep, :]
        stot = np.sum(s ** 2)
        b = dnew
        dnew = min(int(np.sum(np.cumsum(s ** 2 / stot) < (1 - tol))) + 1, dmax)
        if dnew > len(s):
            raise IndexError("compressionSVD!: truncation rank exceeds the number of singular values")
        err_trim[i] = np.sum(s[dnew:] ** 2) / stot
        ent_ent[i] = -2 * np.sum(s ** 2 * np.log(s))
        mps[i] = _r(vh[:dnew, :], dnew, d, b)
        r = u[:, :dnew] * s[:dnew][None, :]
    err_glob = 1 - abs(r.flat[0]) ** 2
    return err_glob, err_trim, ent_ent


# --------------------------------------------------------------------------
# L3 time steppers -- src/timeevolution.jl
# --------------------------------------------------------------------------

def RK4stp_apply_H_half_(dt, H, psi, psi1, psi2, psi3, env1, env2, env3, envop, psi0=None):
    """timeevolution.jl:178-196.

    With ``psi0=None`` this is the 10-argument form (:178-186), whose last
    stage writes the result into ``psi`` while reading it.  With a ``psi0``
    list it is the 11-argument form (:188-196): ``psi`` is left untouched and
    the result goes to ``psi0``.
    """
    out = psi if psi0 is None else psi0
    compress_var_apply_H_left_(psi1, envop, psi, psi, H[0])
    compress_sum_var_apply_H_right_(psi2, [env1], envop, psi1, [1.0, dt / 2], [psi, psi1], H[1])
    compress_sum_var_apply_H_left_(psi3, [env1], envop, psi2, [1.0, dt], [psi, psi2], H[1])
    compress_sum_var_apply_H_right_(out, [env1, env2, env3], envop, psi3,
                                    [-1.0 / 3, 1.0 / 3, 2.0 / 3, 1.0 / 3],
                                    [psi, psi1, psi2, psi3], H[2])


def Arnoldi_expm_(A, rhov, numv, env, envop):
    """timeevolution.jl:88-119: Krylov approximation of exp(A) rhov[0].

    ``env`` is unused, as in the reference.  ``compress_sum_var_apply_H!`` is
    the by-intent restatement above.
    """
    N = np.eye(numv, dtype=CDT)
    Hm = np.zeros((numv, numv), dtype=CDT)
    Hm[0, 0] = scal_op_prod(rhov[0], A, rhov[0])
    for j in range(1, numv):
        coefs = [-Hm[l, j - 1] for l in range(j)] + [1.0]
        vecs = [rhov[l] for l in range(j)] + [rhov[j - 1]]
        compress_sum_var_apply_H_(rhov[j], envop, coefs, vecs, A, 1)
        normalize_lo_(rhov[j])
        Hm[j, j] = scal_op_prod(rhov[j], A, rhov[j])
        for l in range(j):
            Hm[j, l] = scal_op_prod(rhov[j], A, rhov[l])
            Hm[l, j] = scal_op_prod(rhov[l], A, rhov[j])
            N[j, l] = scal_prod(rhov[j], rhov[l])
            N[l, j] = np.conj(N[j, l])
    coef = expm(np.linalg.solve(N, Hm))[:, 0]
    rhov[0] = compress_sum_var(rhov[0], list(coef), rhov, 1)
    normalize_lo_(rhov[0])
    return coef


# --------------------------------------------------------------------------
# steady-state search -- src/steadystate.jl, with the two-operator helpers of
# src/MatrixProductStates.jl:263-312 restated as their @tensor index formulas
# --------------------------------------------------------------------------

def update_lenv_two_op(A0, O1, O2, A, FL):
    """MatrixProductStates.jl:297-301: FL'[a',w1',w2',b'] =
    sum FL[a,w1,w2,b] A[a,s,a'] O1[w1,t,w1',s] O2[w2,u,w2',t] conj(A0[b,u,b'])."""
    T = np.einsum("awxb,asc->wxbsc", FL, A)
    T = np.einsum("wxbsc,wtys->xbcty", T, O1)
    T = np.einsum("xbcty,xuzt->bcyuz", T, O2)
    return np.einsum("bcyuz,bud->cyzd", T, np.conj(A0))


def update_renv_two_op(A0, O1, O2, A, FR):
    """MatrixProductStates.jl:306-310."""
    T = np.einsum("cyzd,asc->yzdas", FR, A)
    T = np.einsum("yzdas,wtys->zdawt", T, O1)
    T = np.einsum("zdawt,xuzt->dawxu", T, O2)
    return np.einsum("dawxu,bud->awxb", T, np.conj(A0))


def local_L_OO_R(FL, O1, O2, FR):
    """The matrix of A_prod_L_OO_R_eigs (MatrixProductStates.jl:258-263) on
    vec(A) (column-major): H[(b,u,d), (a,s,c)]."""
    T = np.einsum("awxb,wtys->axbtys", FL, O1)          # contract w1
    T = np.einsum("axbtys,xuzt->abysuz", T, O2)          # contract w2 and the middle index
    H = np.einsum("abysuz,cyzd->budasc", T, FR)          # contract w1', w2'
    n = H.shape[0] * H.shape[1] * H.shape[2]
    return H.reshape(n, n, order="F")


def _lowest(H):
    """eigs(HM; nev = 1, which = :SR) of a Hermitian local operator, exactly."""
    w, v = np.linalg.eigh(0.5 * (H + H.conj().T))
    return w[0], v[:, 0]


def steady_state_search_(At, L, Lconj, maxsweeps, converge_rat, eigs_tol=None):
    """steadystate.jl:6-86 (in place on the site list At; returns (ae, sweeps)).
    The local eigenproblems are solved exactly (numpy eigh) where the
    reference calls ARPACK to eigs_tol."""
    na = len(At)
    FL = [None] * na
    FR = [None] * na
    FL[0] = np.ones((1, 1, 1, 1), dtype=CDT)
    for j in range(na - 1):
        FL[j + 1] = update_lenv_two_op(At[j], L[j], Lconj[j], At[j], FL[j])
    FR[na - 1] = np.ones((1, 1, 1, 1), dtype=CDT)
    ae = aestart = 0.0
    sw = 0
    for sw in range(1, maxsweeps + 1):
        for j in range(na - 1, 0, -1):
            Dl, d, Dr = At[j].shape
            ae, av = _lowest(local_L_OO_R(FL[j], L[j], Lconj[j], FR[j]))
            if j == na - 1:
                aestart = ae
            Q, R = thin_qr(_r(av, Dl, d * Dr).conj().T)
            At[j] = _r(Q.conj().T, Dl, d, Dr)
            FR[j - 1] = update_renv_two_op(At[j], L[j], Lconj[j], At[j], FR[j])
            At[j - 1] = np.einsum("abl,kl->abk", At[j - 1], R)
        for j in range(na - 1):
            Dl, d, Dr = At[j].shape
            ae, av = _lowest(local_L_OO_R(FL[j], L[j], Lconj[j], FR[j]))
            Q, R = thin_qr(_r(av, Dl * d, Dr))
            At[j] = _r(Q, Dl, d, Dr)
            FL[j + 1] = update_lenv_two_op(At[j], L[j], Lconj[j], At[j], FL[j])
            At[j + 1] = np.einsum("kl,lsb->ksb", R, At[j + 1])
        Dl, d, Dr = At[na - 1].shape
        ae, av = _lowest(local_L_OO_R(FL[na - 1], L[na - 1], Lconj[na - 1], FR[na - 1]))
        At[na - 1] = _r(av, Dl, d, Dr)
        if abs(ae) > converge_rat * abs(aestart) or abs(ae) < 1e-10:
            break
    return ae, sw


# --------------------------------------------------------------------------
# two-site variational compression -- src/compression.jl:110-164, 355-486,
# 574-647, 1072-1273, by intent (the reference's svdecon is never defined; it
# is the thin SVD here, numpy's LAPACK gesdd)
# --------------------------------------------------------------------------

def two_site_sweeps(Ain, coef, A, H, mode, maxsweeps, svd_tol, dlim):
    """Two-site sweeps towards sum_l coef[l] A[l] (+ coef[-1] H A[-1] when H
    is given); mode 0 full sweeps, 1 one left half-sweep, 2 one right
    half-sweep.  Returns (A0, dims).  Environments as in the one-site sums:
    overlap L[bra, ket] / R[ket, bra], operator envs with the conjugated left
    storage; coefficients at the far-end boundary."""
    na = len(Ain)
    nov = len(A) - (1 if H is not None else 0)
    cH = coef[-1] if H is not None else 0.0
    dims = [Ain[0].shape[0]] + [a.shape[2] for a in Ain]
    A0 = [a.copy() for a in Ain]
    E = [[None] * (na + 1) for _ in range(nov)]
    Eop = [None] * (na + 1)

    def project(j):
        T = 0
        for l in range(nov):
            AL = prod_LR_23(E[l][j], A[l][j])
            AR = prod_LR_32(A[l][j + 1], E[l][j + 2])
            T = T + np.einsum("bsc,ctf->bstf", AL, AR)
        X = Y = None
        if H is not None:
            X = prod_AOL(A[nov][j], H[j], Eop[j])
            Y = prod_AOR(A[nov][j + 1], H[j + 1], Eop[j + 2])
            T = T + np.einsum("bscw,cwtf->bstf", X, Y)
        return T, X, Y

    def svd_keep(T):
        b, s, t, f = T.shape
        U, S, Vh = np.linalg.svd(T.reshape(b * s, t * f, order="F"), full_matrices=False)
        keep = S > 0
        U, S, Vh = U[:, keep], S[keep], Vh[keep]
        tot = np.sum(S ** 2)
        k = min(int(np.sum(np.cumsum(S ** 2 / tot) < 1 - svd_tol)) + 1, dlim, len(S))
        return U[:, :k], S[:k], Vh[:k], k

    def left():
        for j in range(na - 2, -1, -1):
            T, X, Y = project(j)
            b, s, t, f = T.shape
            U, S, Vh, k = svd_keep(T)
            dims[j + 1] = k
            A0[j + 1] = Vh.reshape(k, t, f, order="F")
            for l in range(nov):
                E[l][j + 1] = update_renv(A0[j + 1], prod_LR_32(A[l][j + 1], E[l][j + 2]))
            if H is not None:
                Eop[j + 1] = update_re(A0[j + 1], Y)
            if j == 0:
                A0[0] = (U * S).reshape(b, s, k, order="F")

    def right():
        for j in range(na - 1):
            T, X, Y = project(j)
            b, s, t, f = T.shape
            U, S, Vh, k = svd_keep(T)
            dims[j + 1] = k
            A0[j] = U.reshape(b, s, k, order="F")
            for l in range(nov):
                E[l][j + 1] = update_lenv(A0[j], prod_LR_23(E[l][j], A[l][j]))
            if H is not None:
                Eop[j + 1] = update_le(A0[j], X)
            if j == na - 2:
                A0[na - 1] = (S[:, None] * Vh).reshape(k, t, f, order="F")

    one = lambda: np.ones((1, 1), dtype=CDT)           # noqa: E731
    one3 = lambda: np.ones((1, 1, 1), dtype=CDT)       # noqa: E731
    if mode == 2:
        for l in range(nov):
            E[l][na] = one()
            for j in range(na - 1, 1, -1):
                E[l][j] = update_renv_full(Ain[j], A[l][j], E[l][j + 1])
            E[l][0] = coef[l] * one()
        if H is not None:
            Eop[na] = one3()
            for j in range(na - 1, 1, -1):
                Eop[j] = update_re_full(Ain[j], H[j], A[nov][j], Eop[j + 1])
            Eop[0] = cH * one3()
        right()
        return A0, dims
    for l in range(nov):
        E[l][0] = one()
        for j in range(na - 2):
            E[l][j + 1] = update_lenv_full(Ain[j], A[l][j], E[l][j])
        E[l][na] = coef[l] * one()
    if H is not None:
        Eop[0] = one3()
        for j in range(na - 2):
            Eop[j + 1] = update_le_full(Ain[j], H[j], A[nov][j], Eop[j])
        Eop[na] = cH * one3()
    if mode == 1:
        left()
    else:
        for _ in range(maxsweeps):
            left()
            right()
    return A0, dims
