"""GPU parity of the L1 contractions (mps_base.jl) and the device QR against the oracle."""
import numpy as np
import pytest

from oracle import mps_oracle as O
from helpers import crandn, rel

pytestmark = pytest.mark.gpu

SHAPES = [  # Dl, d, Dr, wl, wr, D0 (bond of the second state)
    (1, 2, 2, 1, 4, 2),
    (2, 2, 4, 4, 4, 4),
    (8, 2, 8, 4, 4, 8),
    (16, 9, 16, 12, 12, 16),
    (33, 3, 17, 5, 7, 19),
    (128, 2, 128, 4, 4, 128),
    (64, 9, 64, 14, 14, 64),
]

TOL = 1e-12   # FP64 contractions: relative Frobenius error


@pytest.fixture(scope="module")
def dev():
    import devbuf as D
    return D


@pytest.mark.parametrize("shape", SHAPES)
def test_prod_AOR_update_re(dev, shape, rng):
    Dl, d, Dr, wl, wr, D0 = shape
    print("case", shape, flush=True)
    A = crandn(rng, (Dl, d, Dr))
    Om = crandn(rng, (wl, d, wr, d))
    R = crandn(rng, (Dr, wr, D0))
    ref = O.prod_AOR(A, Om, R)
    got = dev.prod_AOR(A, Om, R)
    assert rel(got, ref) < TOL
    L = crandn(rng, (Dl, wl, D0))
    assert rel(dev.prod_LR(L, ref), O.prod_LR_34(L, ref)) < TOL
    A0 = crandn(rng, (D0, d, D0))
    assert rel(dev.update_re(A0, ref), O.update_re(A0, ref)) < TOL


@pytest.mark.parametrize("shape", SHAPES)
def test_prod_AOL_update_le(dev, shape, rng):
    Dl, d, Dr, wl, wr, D0 = shape
    print("case", shape, flush=True)
    A = crandn(rng, (Dl, d, Dr))
    Om = crandn(rng, (wl, d, wr, d))
    L = crandn(rng, (Dl, wl, D0))
    ref = O.prod_AOL(A, Om, L)
    got = dev.prod_AOL(A, Om, L)
    assert rel(got, ref) < TOL
    R = crandn(rng, (Dr, wr, D0))
    assert rel(dev.prod_LR(ref, R), O.prod_LR_43(ref, R)) < TOL
    A0 = crandn(rng, (D0, d, D0))
    assert rel(dev.update_le(A0, ref), O.update_le(A0, ref)) < TOL


@pytest.mark.parametrize("shape", SHAPES)
def test_overlap_contractions(dev, shape, rng):
    Dl, d, Dr, _, _, D0 = shape
    print("case", shape, flush=True)
    A = crandn(rng, (Dl, d, Dr))
    A0 = crandn(rng, (D0, d, Dr))
    L = crandn(rng, (D0, Dl))
    R = crandn(rng, (Dr, D0))
    assert rel(dev.prod_LR(L, A), O.prod_LR_23(L, A)) < TOL
    assert rel(dev.prod_LR(L.T.copy(), A, co=True), O.prod_LR_co_23(L.T.copy(), A)) < TOL
    assert rel(dev.prod_LR(A, R), O.prod_LR_32(A, R)) < TOL
    AL = O.prod_LR_23(L, A)                 # [D0, d, Dr]
    B0 = crandn(rng, (D0, d, Dr))
    assert rel(dev.update_lenv(B0, AL, 0), O.update_lenv(B0, AL)) < TOL
    assert rel(dev.update_lenv(B0, AL, 1), O.update_lenv_co(B0, AL)) < TOL
    assert rel(dev.update_lenv(B0, AL, 2), O.update_lenv_nc(B0, AL)) < TOL
    AR = O.prod_LR_32(A, R)                 # [Dl, d, D0]
    C0 = crandn(rng, (D0, d, D0))
    assert rel(dev.update_renv(C0, AR), O.update_renv(C0, AR)) < TOL
    Op = crandn(rng, (d, d))
    assert abs(dev.local_exp(A0, A0, Op) - O.local_exp(A0, A0, Op)) < 1e-10 * max(1, abs(O.local_exp(A0, A0, Op)))


@pytest.mark.parametrize("mn", [(8, 4), (64, 32), (256, 128), (1152, 128), (300, 77), (1000, 50),
                                (2304, 256), (2600, 40), (1152, 8), (1280, 256), (900, 3)])
@pytest.mark.parametrize("view", [False, True])
def test_qrq_full_rank(dev, mn, view, rng):
    m, n = mn
    print("case", mn, view, flush=True)
    M = crandn(rng, (m, n))
    Q = dev.qrq(M, transposed_view=view)
    Qr = O.qrq(M)
    assert np.abs(Q - Qr).max() < 1e-12
    Qd, R = dev.qrq(M, want_r=True)
    assert np.abs(Qd @ R - M).max() < 1e-12 * np.abs(M).max() * n


@pytest.mark.parametrize("mnr", [(1152, 128, 128), (1000, 50, 50), (1152, 128, 20), (2304, 256, 256),
                                 (2304, 256, 30), (1280, 256, 256), (512, 256, 256), (512, 256, 40),
                                 (300, 77, 77), (768, 256, 256), (700, 100, 30), (288, 32, 32),
                                 (256, 128, 128), (256, 128, 20), (200, 40, 40), (64, 32, 32), (64, 17, 5)])
def test_qrq_single_launch_matches_per_panel(dev, mnr, rng):
    # the single-launch (chain) factorization and the launch-per-panel one
    # build the same zgeqr2 reflectors: Q and R agree to rounding, with and
    # without the rank cutoff (rank-r panels plus rounding-level noise)
    m, n, r = mnr
    print("case", mnr, flush=True)
    M = crandn(rng, (m, r)) @ crandn(rng, (r, n))
    if r < n:
        M = M + 1e-17 * crandn(rng, (m, n))
    from mpsamd import default_context
    ctx = default_context()
    tol = 1e-13 if r < n else 0.0
    ctx.set_qr_rank_tol(tol)
    try:
        ctx.set_qr_single_launch(True)
        Qc, Rc = dev.qrq(M, want_r=True)
        ctx.set_qr_single_launch(False)
        Qp, Rp = dev.qrq(M, want_r=True)
    finally:
        ctx.set_qr_single_launch(None)
        ctx.set_qr_rank_tol(0.0)
    scale = np.abs(M).max() * n
    assert np.abs(Qc - Qp).max() < 1e-12
    assert np.abs(Rc - Rp).max() < 1e-12 * scale
    assert np.abs(Qc.conj().T @ Qc - np.eye(n)).max() < 1e-12


def test_qrq_zero_padded_closed_form(dev):
    # zero-padded product-state panel: LAPACK gives tau = 0 for the zero
    # columns; the completion is fixed by the one non-trivial reflector.
    M = np.zeros((8, 4), dtype=np.complex128)
    M[3, 0] = 1.0
    M[5, 2] = 0.5
    Q = dev.qrq(M)
    assert np.array_equal(Q, O.qrq(M))


@pytest.mark.parametrize("tol", [1e-13, 1e-8])
@pytest.mark.parametrize("mnr", [(64, 16, 3), (600, 70, 9), (1152, 128, 20), (2304, 40, 5), (1100, 256, 60),
                                 (1152, 128, 125)])
def test_qrq_rank_cutoff(dev, tol, mnr, rng):
    # low-rank panel plus rounding-level noise: with the cutoff both sides take
    # the identity reflector on the noise columns, so the completion agrees.
    m, n, r = mnr
    M = crandn(rng, (m, r)) @ crandn(rng, (r, n)) + 1e-17 * crandn(rng, (m, n))
    from mpsamd import default_context
    ctx = default_context()
    ctx.set_qr_rank_tol(tol)
    try:
        Q = dev.qrq(M)
    finally:
        ctx.set_qr_rank_tol(0.0)
    O.set_qr_rank_tol(tol)
    Qr = O.qrq(M)
    assert np.abs(Q - Qr).max() < 1e-10
    assert np.abs(Q.conj().T @ Q - np.eye(n)).max() < 1e-12
