"""Pin the oracle's QR (orthonormalisation.jl) against LAPACK invariants and closed forms."""
import numpy as np
import pytest

from oracle import mps_oracle as O
from helpers import crandn


@pytest.mark.parametrize("mn", [(6, 3), (32, 32), (128, 40)])
def test_qrq_invariants(mn, rng):
    m, n = mn
    A = crandn(rng, (m, n))
    Q = O.qrq(A)
    assert np.abs(Q.conj().T @ Q - np.eye(n)).max() < 1e-13
    R = Q.conj().T @ A
    assert np.abs(np.tril(R, -1)).max() < 1e-12 * np.abs(A).max()
    # LAPACK convention: R diagonal real, beta = -sign(Re alpha) ||x||
    assert np.abs(np.diag(R).imag).max() < 1e-12
    assert np.sign(np.diag(R).real[0]) == -np.sign(A[0, 0].real)


def test_householder_restatement_equals_lapack(rng):
    for m, n in [(10, 4), (64, 64), (200, 64)]:
        A = crandn(rng, (m, n))
        qr, tau = O.householder_geqr2(A)
        Q = O.householder_ung2r(qr, tau)
        assert np.abs(Q - O.qrq(A)).max() < 1e-13
        ql, tl = O.geqrf(A)
        assert np.abs(tau - tl).max() < 1e-13


def test_zero_padded_closed_form():
    # mpsproductstate panel with state [0, 1] on a 2-dim left bond: column 0
    # is e_{Dl}; every other column is zero -> tau = 0 there (H = I), and Q
    # is fixed by the one reflector: Q[:,0] = -e_2, Q[:,2] = -e_0, else e_k.
    M = np.zeros((4, 3), dtype=complex)
    M[2, 0] = 1.0
    Q = O.qrq(M)
    ref = np.zeros((4, 3), dtype=complex)
    ref[2, 0] = -1.0
    ref[1, 1] = 1.0
    ref[0, 2] = -1.0
    assert np.array_equal(Q, ref)
    qr, tau = O.householder_geqr2(M)
    assert tau[1] == 0 and tau[2] == 0
    assert np.array_equal(O.householder_ung2r(qr, tau), ref)


def test_rank_cutoff_identity_branch(rng):
    m, n = 40, 12
    A = crandn(rng, (m, 2)) @ crandn(rng, (2, n))
    noisy = A + 1e-17 * crandn(rng, (m, n))
    qr, tau = O.householder_geqr2(noisy, tol=1e-13)
    assert np.count_nonzero(tau) == 2          # the rank, not the noise
    Q = O.householder_ung2r(qr, tau)
    assert np.abs(Q.conj().T @ Q - np.eye(n)).max() < 1e-13
    # the span of the first two columns is the column space of A
    P = Q[:, :2] @ Q[:, :2].conj().T
    assert np.abs(P @ A - A).max() < 1e-12 * np.abs(A).max()


def test_leftorth_rightorth(rng):
    A = crandn(rng, (3, 2, 5))
    AL, C = O.leftorth(A)
    assert np.allclose(O.prod_LR_32(AL, C), A)
    AR, C = O.rightorth(A)
    assert np.allclose(O.prod_LR_23(C, AR), A)
    Q = O.rightorth_(A.copy())
    M = Q.reshape(3, 10, order="F")
    assert np.abs(M @ M.conj().T - np.eye(3)).max() < 1e-13
