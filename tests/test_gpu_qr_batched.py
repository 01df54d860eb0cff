"""leftorth! (orthonormalisation.jl:164-171) on every copy of a lockstep batch
at once, with the single-launch factorization (each copy's panels on one XCD
of its own) against the launch-per-panel one and against the same copy on a
plain context.  Copies of one batch have different ranks, so their identity
reflectors (rank cutoff 1e-13) differ from copy to copy."""
import numpy as np
import pytest

import mpsamd as M
from helpers import crandn
from mpsamd.trajectory import _state_bytes

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _sites(rng, d, D, nb):
    # copy b: site 0 of rank r_b (full rank for b % 3 == 0), site 1 random
    out = []
    m = D * d
    for b in range(nb):
        r = D if b % 3 == 0 else max(1, (D * (b % 3)) // 5 + b)
        A = crandn(rng, (m, r)) @ crandn(rng, (r, D))
        if r < D:
            A = A + 1e-17 * crandn(rng, (m, D))
        # site (D, d, D) whose (D*d) x D matricisation is A (left index fastest)
        s0 = np.asfortranarray(A.reshape((D, d, D), order="F"))
        s1 = crandn(rng, (D, d, D))
        out.append((s0, s1))
    return out


def _leftorth_batched(sites, d, D, single_launch):
    nb = len(sites)
    dims = [D, D, D]
    ctx = M.Context.batched(nb, _state_bytes(dims, d) + 4096, M.batch_workspace_bytes(nb, D, d, 1), 0)
    comp = ctx.companion()
    ctx.set_qr_rank_tol(TOL)
    ctx.set_qr_single_launch(single_launch)
    g = M.MPS(2, d, dims, ctx)
    views = [g.view(b, comp) for b in range(nb)]
    for v, s in zip(views, sites):
        v.upload(list(s))
    M.leftorth_(g, 0)
    ctx.synchronize()
    out = [v.to_host() for v in views]
    for v in views:
        v.__del__()
    g.__del__()
    comp.close()
    ctx.close()
    return out


def _leftorth_plain(s, d, D):
    ctx = M.Context(0)
    ctx.set_qr_rank_tol(TOL)
    ctx.set_qr_single_launch(True)
    st = M.MPS(2, d, [D, D, D], ctx)
    st.upload(list(s))
    M.leftorth_(st, 0)
    ctx.synchronize()
    out = st.to_host()
    st.__del__()
    ctx.close()
    return out


@pytest.mark.parametrize("d,D,nb", [(2, 256, 16), (3, 256, 9), (9, 128, 8), (2, 150, 3), (9, 100, 10),
                                     (2, 128, 16), (2, 32, 8), (9, 24, 5)])
def test_batched_leftorth_single_launch(d, D, nb, rng):
    print("case d=%d D=%d nb=%d" % (d, D, nb), flush=True)
    sites = _sites(rng, d, D, nb)
    chain = _leftorth_batched(sites, d, D, True)
    panel = _leftorth_batched(sites, d, D, False)
    for b in range(nb):
        q_c, r_c = chain[b]
        q_p, r_p = panel[b]
        scale = np.abs(sites[b][0]).max() * D
        assert np.isfinite(q_c).all() and np.isfinite(r_c).all(), b
        assert np.abs(q_c - q_p).max() < 1e-12, b
        # site 1 holds R times the old site 1: compare relative to its size
        assert np.abs(r_c - r_p).max() < 1e-12 * scale * np.abs(sites[b][1]).max() * D, b
        Q = q_c.reshape((D * d, D), order="F")
        assert np.abs(Q.conj().T @ Q - np.eye(D)).max() < 1e-12, b
    # the first, a rank-deficient and the last copy against the same site on a plain context
    for b in sorted({0, 1, nb - 1}):
        q1, r1 = _leftorth_plain(sites[b], d, D)
        assert np.abs(chain[b][0] - q1).max() < 1e-12, b
        assert np.abs(chain[b][1] - r1).max() < 1e-12 * np.abs(r1).max() * D, b
