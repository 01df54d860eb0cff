"""The headline workload at its own size against the oracle: one RK4 step
(RK4stp_apply_H_half!, 10-argument form) of the Rydberg-ladder density matrix
at N = 100, d = 9, n_exp = 4 (w = 14), D_max = 128, dt = 0.005, positions
from seed 1, started from mpsgroundstate, on the device and in the FP64 CPU
oracle.  The ground state is a product state zero-padded to D = 128, so both
sides use the QR rank cutoff 1e-13: with the plain zlarfg rule the padded
directions are completed from rounding noise, which differs between any two
implementations.  Even with the cutoff, columns near its threshold may be
completed differently, so the two results can differ by a gauge transformation
site by site (measured: up to 0.29 relative on single sites) while
representing the same state; the checks are therefore on the state: fidelity,
trace and populations.  The oracle step takes about 20 s of host time.
"""
import numpy as np
import pytest

import mpsamd as M
from mpsamd import models
from oracle import mps_oracle as O
from helpers import copy_sites

pytestmark = pytest.mark.gpu


def test_rydberg_density_matrix_N100_D128_one_step():
    N, D, dt, n_exp, seed, tol = 100, 128, 0.005, 4, 1, 1e-13
    p = models.RydbergParams(n_exp=n_exp)
    z = models.positions(N, float(N), seed)
    L = models.rydberg_rk4_mpos(p, z, dt)
    wop = [1] + [p.w] * (N - 1) + [1]
    ctx = M.default_context()
    ctx.set_qr_rank_tol(tol)
    O.set_qr_rank_tol(tol)
    try:
        print("case N=100 D=128 d=9 w=14: device step", flush=True)
        rho, dims = M.mpsgroundstate(N, D, 9)
        r = [rho.similar() for _ in range(3)]
        for s in r:
            s.copy_from(rho)
        e = [M.build_env(dims, dims) for _ in range(3)]
        eop = M.build_env(dims, dims, wop)
        gL = [M.MPO(h) for h in L]
        M.RK4stp_apply_H_half_(dt, gL, rho, *r, *e, eop)
        got = rho.to_host()
        idm = M.MPS.from_host(models.identity_vec_mps(N))
        nem = M.MPS.from_host(models.sum_local_vec_mps(N, models.HEE3))
        nrm = M.MPS.from_host(models.sum_local_vec_mps(N, models.HSS3))
        tr_d = M.scal_prod_no_conj(idm, rho)
        ne_d = M.scal_prod_no_conj(nem, rho) / tr_d
        nr_d = M.scal_prod_no_conj(nrm, rho) / tr_d

        print("case N=100 D=128 d=9 w=14: oracle step", flush=True)
        ref, rdims = O.mpsgroundstate(N, D, 9)
        assert list(rdims) == list(dims)
        st = [copy_sites(ref) for _ in range(3)]
        envs = [O.build_env(dims, dims) for _ in range(3)]
        O.RK4stp_apply_H_half_(dt, L, ref, *st, *envs, O.build_env_op(dims, dims, wop))
    finally:
        ctx.set_qr_rank_tol(0.0)
        O.set_qr_rank_tol(0.0)
    tr_o = O.scal_prod_no_conj(models.identity_vec_mps(N), ref)
    ne_o = O.scal_prod_no_conj(models.sum_local_vec_mps(N, models.HEE3), ref) / tr_o
    nr_o = O.scal_prod_no_conj(models.sum_local_vec_mps(N, models.HSS3), ref) / tr_o
    ab, aa, bb = O.scal_prod(got, ref), O.scal_prod(got, got), O.scal_prod(ref, ref)
    infid = 1.0 - abs(ab) ** 2 / (abs(aa) * abs(bb))
    print("trace", tr_d, tr_o, "n_e", ne_d, ne_o, "n_r", nr_d, nr_o, "1 - fidelity", infid, flush=True)
    assert abs(ne_o) > 1e-6   # the step did excite the chain
    assert abs(tr_d - tr_o) < 1e-10
    assert abs(ne_d - ne_o) < 1e-10 * abs(ne_o)
    assert abs(nr_d - nr_o) < 1e-10 * max(abs(nr_o), 1e-12) + 1e-20
    assert infid < 1e-12
