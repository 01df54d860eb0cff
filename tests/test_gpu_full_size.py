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


def _fixture():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rydberg_N100_D128_trace.npz")
    return np.load(path)


def test_rydberg_density_matrix_N100_D128_trajectory_against_oracle_fixture():
    # The oracle's per-step trace, populations, <rho|rho> and three fingerprints
    # <X_k|rho> (tests/golden/make_full_size_trace.py) over its recorded steps,
    # against the device from the same start, at the cutoff both sides use.
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_full_size_trace import fingerprint_mps
    f = _fixture()
    N, D, dt, n_exp, seed = 100, 128, float(f["dt"]), 4, 1
    p = models.RydbergParams(n_exp=n_exp)
    assert np.allclose(p.u, f["u"], rtol=1e-12) and np.allclose(p.kappa, f["kappa"], rtol=1e-12)
    z = models.positions(N, float(N), seed)
    wop = [1] + [p.w] * (N - 1) + [1]
    ctx = M.Context(0)
    ctx.set_qr_rank_tol(float(f["tol"]))
    rho, dims = M.mpsgroundstate(N, D, 9, ctx)
    r = [rho.similar() for _ in range(3)]
    for s in r:
        s.copy_from(rho)
    e = [M.build_env(dims, dims, ctx=ctx) for _ in range(3)]
    eop = M.build_env(dims, dims, wop, ctx=ctx)
    gL = [M.MPO(h, ctx) for h in models.rydberg_rk4_mpos(p, z, dt)]
    idm = M.MPS.from_host(models.identity_vec_mps(N), ctx)
    nem = M.MPS.from_host(models.sum_local_vec_mps(N, models.HEE3), ctx)
    nrm = M.MPS.from_host(models.sum_local_vec_mps(N, models.HSS3), ctx)
    X = [M.MPS.from_host(fingerprint_mps(k), ctx) for k in range(3)]
    worst = {"trace": 0.0, "n_e": 0.0, "n_r": 0.0, "hs": 0.0, "fp": 0.0}
    # <n_r> is real in exact arithmetic and tiny here (4e-11 .. 6e-6): the
    # imaginary part of the oracle's value is its own rounding error, so the
    # relative 1e-9 bound on it holds above that floor
    nr_floor = 2.0 * float(np.abs(np.imag(f["n_r"])).max())
    for s, _ in enumerate(f["steps"]):
        M.RK4stp_apply_H_half_(dt, gL, rho, *r, *e, eop)
        tr = M.scal_prod_no_conj(idm, rho)
        got = {"trace": tr, "n_e": M.scal_prod_no_conj(nem, rho) / tr, "n_r": M.scal_prod_no_conj(nrm, rho) / tr,
               "hs": M.scal_prod(rho, rho)}
        for k in ("n_e", "hs"):
            worst[k] = max(worst[k], abs(got[k] - f[k][s]) / abs(f[k][s]))
        worst["n_r"] = max(worst["n_r"], max(abs(got["n_r"] - f["n_r"][s]) - nr_floor, 0.0) / abs(f["n_r"][s]))
        worst["trace"] = max(worst["trace"], abs(got["trace"] - f["trace"][s]))
        fp = np.array([M.scal_prod(x, rho) for x in X])
        worst["fp"] = max(worst["fp"], float(np.max(np.abs(fp - f["fp"][s]) / np.abs(f["fp"][s]))))
        print("step", s + 1, {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    assert worst["trace"] < 1e-10 and worst["n_e"] < 1e-10 and worst["hs"] < 1e-10
    assert worst["n_r"] < 1e-9 and worst["fp"] < 1e-9


def test_rydberg_density_matrix_N100_D128_200_steps_properties():
    # cfg1's 200 steps at the full size: no divergence, the trace stays 1 up
    # to the bond-128 truncation (measured: 1.2e-5 short after 200 steps),
    # populations stay in [0, 1] per atom and rise monotonically under the
    # weak probe from the ground state, and purity <rho|rho> stays <= 1.
    N, D, dt, n_exp, seed = 100, 128, 0.005, 4, 1
    p = models.RydbergParams(n_exp=n_exp)
    z = models.positions(N, float(N), seed)
    wop = [1] + [p.w] * (N - 1) + [1]
    ctx = M.Context(0)
    ctx.set_qr_rank_tol(1e-13)
    rho, dims = M.mpsgroundstate(N, D, 9, ctx)
    r = [rho.similar() for _ in range(3)]
    for s in r:
        s.copy_from(rho)
    e = [M.build_env(dims, dims, ctx=ctx) for _ in range(3)]
    eop = M.build_env(dims, dims, wop, ctx=ctx)
    gL = [M.MPO(h, ctx) for h in models.rydberg_rk4_mpos(p, z, dt)]
    idm = M.MPS.from_host(models.identity_vec_mps(N), ctx)
    nem = M.MPS.from_host(models.sum_local_vec_mps(N, models.HEE3), ctx)
    nrm = M.MPS.from_host(models.sum_local_vec_mps(N, models.HSS3), ctx)
    prev = 0.0
    for s in range(1, 201):
        M.RK4stp_apply_H_half_(dt, gL, rho, *r, *e, eop)
        if s % 20 == 0:
            tr = M.scal_prod_no_conj(idm, rho)
            ne = (M.scal_prod_no_conj(nem, rho) / tr)
            nr = (M.scal_prod_no_conj(nrm, rho) / tr)
            hs = M.scal_prod(rho, rho).real
            print("step", s, "trace", tr, "n_e", ne, "n_r", nr, "purity", hs, flush=True)
            assert abs(tr.real - 1.0) < 1e-4 and abs(tr.imag) < 1e-6
            assert 0.0 < ne.real < N and 0.0 < nr.real < N and abs(ne.imag) < 1e-6
            assert ne.real > prev
            assert hs <= 1.0 + 1e-10
            prev = ne.real
    pops = [M.scal_prod_no_conj(M.MPS.from_host(models.local_population_vec_mps(N, j, models.HEE3), ctx), rho).real
            for j in range(N)]
    assert min(pops) >= -1e-8 and max(pops) <= 1.0


@pytest.mark.parametrize("D,steps", [(128, 10), (256, 4)])
def test_rydberg_density_matrix_N100_precision_forms(D, steps):
    # The precision form at the full size (D_max 128 and 256): RK4 steps with
    # the operator contractions in complex FP32, without and with 2 FP64
    # refinement half-sweeps per call, against the FP64 run from the same
    # start.  Populations within 1e-6, 1 - fidelity below 1e-12, and the FP32
    # kernels visibly ran (the populations differ by more than FP64 rounding).
    print("case D", D, "steps", steps, flush=True)
    N, dt, n_exp, seed = 100, 0.005, 4, 1
    p = models.RydbergParams(n_exp=n_exp)
    z = models.positions(N, float(N), seed)
    wop = [1] + [p.w] * (N - 1) + [1]
    ctx = M.Context(0)
    ctx.set_qr_rank_tol(1e-13)
    rho, dims = M.mpsgroundstate(N, D, 9, ctx)
    rho0 = rho.to_host()
    r = [rho.similar() for _ in range(3)]
    e = [M.build_env(dims, dims, ctx=ctx) for _ in range(3)]
    eop = M.build_env(dims, dims, wop, ctx=ctx)
    gL = [M.MPO(h, ctx) for h in models.rydberg_rk4_mpos(p, z, dt)]
    idm = M.MPS.from_host(models.identity_vec_mps(N), ctx)
    nem = M.MPS.from_host(models.sum_local_vec_mps(N, models.HEE3), ctx)
    nrm = M.MPS.from_host(models.sum_local_vec_mps(N, models.HSS3), ctx)

    def run(prec, nref):
        rho.upload(rho0)
        for s in r:
            s.copy_from(rho)
        ctx.set_precision(prec)
        ctx.set_refine_sweeps(nref)
        try:
            for _ in range(steps):
                M.RK4stp_apply_H_half_(dt, gL, rho, *r, *e, eop)
        finally:
            ctx.set_precision("fp64")
            ctx.set_refine_sweeps(0)
        return rho.similar().copy_from(rho)

    def pops(x):
        tr = M.scal_prod_no_conj(idm, x)
        return (M.scal_prod_no_conj(nem, x) / tr).real, (M.scal_prod_no_conj(nrm, x) / tr).real

    def infidelity(a, b):
        ab = abs(M.scal_prod(a, b))
        return 1.0 - ab * ab / (abs(M.scal_prod(a, a)) * abs(M.scal_prod(b, b)))

    ref = run("fp64", 0)
    ne64, nr64 = pops(ref)
    assert ne64 > 1e-4   # the steps did excite the chain
    for nref in (0, 2):
        got = run("fp32", nref)
        ne, nr = pops(got)
        inf = infidelity(got, ref)
        print("refine", nref, "n_e", ne, ne64, "n_r", nr, nr64, "1 - fidelity", inf, flush=True)
        assert abs(ne - ne64) < 1e-6 and abs(nr - nr64) < 1e-6
        assert inf < 1e-12
        assert abs(ne - ne64) > 1e-14
