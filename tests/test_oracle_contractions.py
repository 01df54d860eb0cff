"""Pin the oracle's L1 contractions (mps_base.jl) against brute-force einsum of their index formulas.

The einsum strings are the mathematical definitions (independent of the
reshape/permutedims chains the oracle restates), so a dropped conjugate, a
transposed operand or a wrong permutation fails here.
"""
import numpy as np
import pytest

from oracle import mps_oracle as O
from helpers import crandn

SHAPES = [(1, 2, 3, 1, 4, 2), (3, 2, 4, 4, 4, 5), (4, 3, 2, 5, 3, 3), (6, 9, 5, 6, 6, 4)]


@pytest.mark.parametrize("s", SHAPES)
def test_prod_AOR(s, rng):
    Dl, d, Dr, wl, wr, D0 = s
    A, Om, R = crandn(rng, (Dl, d, Dr)), crandn(rng, (wl, d, wr, d)), crandn(rng, (Dr, wr, D0))
    ref = np.einsum("asx,xvc,wtvs->awtc", A, R, Om)
    assert np.allclose(O.prod_AOR(A, Om, R), ref, atol=1e-13)


@pytest.mark.parametrize("s", SHAPES)
def test_prod_AOL(s, rng):
    Dl, d, Dr, wl, wr, D0 = s
    A, Om, L = crandn(rng, (Dl, d, Dr)), crandn(rng, (wl, d, wr, d)), crandn(rng, (Dl, wl, D0))
    ref = np.einsum("awb,asy,wtvs->btyv", L.conj(), A, Om)
    assert np.allclose(O.prod_AOL(A, Om, L), ref, atol=1e-13)


@pytest.mark.parametrize("s", SHAPES)
def test_prod_LR_and_env_updates(s, rng):
    Dl, d, Dr, wl, wr, D0 = s
    L3 = crandn(rng, (Dl, wl, D0))
    R4 = crandn(rng, (Dl, wl, d, Dr))
    assert np.allclose(O.prod_LR_34(L3, R4), np.einsum("awb,awsc->bsc", L3.conj(), R4), atol=1e-13)
    L4 = crandn(rng, (D0, d, Dr, wr))
    R3 = crandn(rng, (Dr, wr, Dl))
    assert np.allclose(O.prod_LR_43(L4, R3), np.einsum("bsyv,yvc->bsc", L4, R3), atol=1e-13)
    L2 = crandn(rng, (D0, Dl))
    A = crandn(rng, (Dl, d, Dr))
    assert np.allclose(O.prod_LR_23(L2, A), np.einsum("ba,asc->bsc", L2, A), atol=1e-13)
    assert np.allclose(O.prod_LR_co_23(L2.T, A), np.einsum("ab,asc->bsc", L2.T, A), atol=1e-13)
    R2 = crandn(rng, (Dr, D0))
    assert np.allclose(O.prod_LR_32(A, R2), np.einsum("asx,xc->asc", A, R2), atol=1e-13)
    # update_re: R'[a,w,b] = sum AOR[a,w,s,c] conj(A0[b,s,c])
    AOR = crandn(rng, (Dl, wl, d, Dr))
    A0 = crandn(rng, (D0, d, Dr))
    assert np.allclose(O.update_re(A0, AOR), np.einsum("awsc,bsc->awb", AOR, A0.conj()), atol=1e-13)
    # update_le: L'[y,v,c] = sum conj(AOL[b,s,y,v]) A0[b,s,c]
    AOL = crandn(rng, (D0, d, Dr, wr))
    B0 = crandn(rng, (D0, d, Dl))
    assert np.allclose(O.update_le(B0, AOL), np.einsum("bsyv,bsc->yvc", AOL.conj(), B0), atol=1e-13)
    AR = crandn(rng, (Dl, d, Dr))
    C0 = crandn(rng, (D0, d, Dr))
    assert np.allclose(O.update_renv(C0, AR), np.einsum("asc,bsc->ab", AR, C0.conj()), atol=1e-13)
    AL = crandn(rng, (D0, d, Dl))
    assert np.allclose(O.update_lenv(C0, AL), np.einsum("bsc,bsa->ca", C0.conj(), AL), atol=1e-13)
    assert np.allclose(O.update_lenv_co(C0, AL), np.einsum("bsa,bsc->ac", AL, C0.conj()), atol=1e-13)
    assert np.allclose(O.update_lenv_nc(C0, AL), np.einsum("bsc,bsa->ca", C0, AL), atol=1e-13)


def test_local_exp(rng):
    A = crandn(rng, (3, 4, 5))
    A0 = crandn(rng, (3, 4, 5))
    Op = crandn(rng, (4, 4))
    ref = np.einsum("ts,asb,atb->", Op, A, A0.conj())
    assert np.isclose(O.local_exp(A0, A, Op), ref, atol=1e-12)


def test_mpo_products(rng):
    m1 = crandn(rng, (2, 3, 2, 3))
    m2 = crandn(rng, (3, 3, 1, 3))
    got = O.apply_site_MPOtoMPO(m1, m2)
    # operator product m2*m1 with combined bonds (w1 fastest)
    ref = np.einsum("ask,bktc->abscd t".replace(" ", ""), m2, m1) if False else None
    T = np.einsum("xsyk,wkvt->wxsvyt", m2, m1)      # [w1, w2, s, v1, v2, t]
    ref = T.reshape(2 * 3, 3, 2 * 1, 3, order="F")
    assert np.allclose(got, ref, atol=1e-13)
    ms = crandn(rng, (2, 3, 4))
    got = O.apply_site_MPOtoMPS(ms, m1)
    ref = np.einsum("atb,lsrt->alsbr", ms, m1).reshape(4, 3, 8, order="F")
    assert np.allclose(got, ref)
    c = O.conj_site_mpo(m1)
    assert np.allclose(c, np.transpose(m1.conj(), (0, 3, 2, 1)))


def test_apply_site_operator(rng):
    A = crandn(rng, (3, 4, 2))
    op = crandn(rng, (4, 4))
    ref = np.einsum("st,atb->asb", op, A)
    got = O.apply_site_operator_(A.copy(), op)
    assert np.allclose(got, ref, atol=1e-13)
