"""The oracle's two-site sweeps (compression.jl:110-164, 355-486, 574-647,
1072-1273, by intent with the thin SVD for svdecon) against dense linear
algebra: exact targets at full bond dimension for every mode, the truncation
rule on the first bond of a left half-sweep against the target's dense
Schmidt values, and a product target collapsing every bond to 1."""
import numpy as np
import pytest

from oracle import dense as Dn
from oracle import mps_oracle as O
from helpers import random_mpo


def _setup(rng, N=5, d=2, w=3):
    H = random_mpo(rng, N, d, w)
    A, _ = O.mpsrandom(N, 8, d, rng)
    B, _ = O.mpsrandom(N, 8, d, rng)
    Ain, _ = O.mpsrandom(N, 8, d, rng)
    return H, A, B, Ain


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_two_site_exact_at_full_bond(mode, rng):
    H, A, B, Ain = _setup(rng)
    if mode == 2:
        O.rightorthonormalizationQR_(Ain)
    coef = [0.3 - 0.1j, 1.2 + 0.5j]
    A0, dims = O.two_site_sweeps(Ain, coef, [B, A], H, mode, 2, 0.0, 64)
    # a right-first sweep injects the operator term's coefficient through the
    # conjugated left operator environment, so it acts as conj(c), as the
    # reference's right sweeps do (compression.jl:1031 with prod_AOL's L^H)
    cH = np.conj(coef[1]) if mode == 2 else coef[1]
    want = coef[0] * Dn.mps_to_dense(B) + cH * (Dn.mpo_to_dense(H) @ Dn.mps_to_dense(A))
    got = Dn.mps_to_dense(A0)
    assert np.linalg.norm(got - want) < 1e-12 * np.linalg.norm(want)
    assert dims == [1, 2, 4, 4, 2, 1]


def test_two_site_truncation_rule_first_bond(rng):
    H, A, B, Ain = _setup(rng)
    target = Dn.mpo_to_dense(H) @ Dn.mps_to_dense(A)
    N = 5
    # Schmidt values across the cut before the last site
    S = np.linalg.svd(target.reshape(2 ** (N - 1), 2, order="F"), compute_uv=False)
    S = S[S > 0]
    w = np.cumsum(S ** 2 / np.sum(S ** 2))
    tol = 0.5 * (1 - w[0])                       # keeps exactly two values by the rule
    want = int(np.sum(w < 1 - tol)) + 1
    A0, dims = O.two_site_sweeps(Ain, [1.0], [A], H, 1, 1, tol, 64)
    assert dims[N - 1] == want == 2


def test_two_site_product_target_collapses_bonds(rng):
    N, d = 6, 2
    v = [rng.standard_normal(d) + 1j * rng.standard_normal(d) for _ in range(N)]
    prod = [x.reshape(1, d, 1).astype(complex) for x in v]
    Ain, _ = O.mpsrandom(N, 8, d, rng)
    A0, dims = O.two_site_sweeps(Ain, [1.0], [prod], None, 0, 1, 1e-12, 8)
    assert dims == [1] * (N + 1)
    want = Dn.mps_to_dense(prod)
    assert np.linalg.norm(Dn.mps_to_dense(A0) - want) < 1e-12 * np.linalg.norm(want)
