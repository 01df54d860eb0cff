"""Two-site sweeps with SVD truncation on the device (mps_compress_two_site;
compression.jl:110-164, 355-486, 574-647, 1072-1273 by intent) against the
oracle's: the same bonds after truncation and the same state (dense vectors,
1e-10), for every mode, overlap-only and operator targets, and the RK4 half
step built from two-site sweeps (timeevolution.jl:208-216)."""
import numpy as np
import pytest

import mpsamd as M
from oracle import dense as Dn
from oracle import mps_oracle as O
from helpers import random_mpo, rel
from mpsamd import models

pytestmark = pytest.mark.gpu


def _dev(s):
    return M.MPS.from_host(s)


@pytest.mark.parametrize("mode,with_H", [(0, True), (1, True), (2, True), (0, False)])
def test_two_site_against_oracle(mode, with_H, rng):
    N, d, w, dlim, tol = 6, 2, 3, 5, 1e-8
    print("case two-site mode", mode, "H", with_H, flush=True)
    H = random_mpo(rng, N, d, w)
    A, _ = O.mpsrandom(N, 8, d, rng)
    B, _ = O.mpsrandom(N, 8, d, rng)
    Ain, _ = O.mpsrandom(N, 4, d, rng)
    if mode == 2:
        O.rightorthonormalizationQR_(Ain)
    coef = [0.7, 0.9 + 0.2j] if with_H else [0.7, 0.4 - 0.3j]
    terms = [B, A]
    ref, rdims = O.two_site_sweeps(Ain, coef, terms, H if with_H else None, mode, 2, tol, dlim)
    dims = []
    fn = {0: M.compress_sum_var_apply_H_two_site, 1: M.compress_sum_var_apply_H_two_site_left,
          2: M.compress_sum_var_apply_H_two_site_right}[mode]
    if with_H:
        g = fn(_dev(Ain), coef, [_dev(B), _dev(A)], M.MPO(H), 2, tol, dims, dlim)
    else:
        g = M.compress_sum_var_two_site(_dev(Ain), coef, [_dev(B), _dev(A)], 2, tol, dims, dlim)
    assert dims == rdims and max(dims) <= dlim
    assert rel(Dn.mps_to_dense(g.to_host()), Dn.mps_to_dense(ref)) < 1e-10


def test_rk4_full_two_site_against_oracle():
    N, D, dt, dlim, tol = 6, 8, 0.05, 8, 0.0      # full bond dimension: exact stages
    z = models.positions(N, float(N), 2)
    H = models.spin_rk4_mpos(z, 0.0, dt)
    psi, dims = O.mpsrandom(N, D, 2, np.random.default_rng(9))
    c4 = [-1.0 / 3, 1.0 / 3, 2.0 / 3, 1.0 / 3]
    # the oracle's composition of the same two-site sweeps
    p1, _ = O.two_site_sweeps(psi, [1.0], [psi], H[0], 1, 1, tol, dlim)
    p2, _ = O.two_site_sweeps(p1, [1.0, dt / 2], [psi, p1], H[1], 2, 1, tol, dlim)
    p3, _ = O.two_site_sweeps(p2, [1.0, dt], [psi, p2], H[1], 1, 1, tol, dlim)
    ref, rdims = O.two_site_sweeps(p3, c4, [psi, p1, p2, p3], H[2], 2, 1, tol, dlim)
    dd = []
    g = M.RK4stp_apply_H_half_full_two_site(dt, [M.MPO(h) for h in H], _dev(psi), tol, dd, dlim)
    assert dd == rdims
    assert rel(Dn.mps_to_dense(g.to_host()), Dn.mps_to_dense(ref)) < 1e-10
    # and against dense RK4 (every two-site stage is exact at full bond dimension)
    Hd = (Dn.mpo_to_dense(H[1]))
    x = Dn.mps_to_dense(psi)
    k1 = Hd @ x
    k2 = Hd @ (x + dt / 2 * k1)
    k3 = Hd @ (x + dt / 2 * k2)
    k4 = Hd @ (x + dt * k3)
    want = x + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    assert rel(Dn.mps_to_dense(g.to_host()), want) < 1e-10
