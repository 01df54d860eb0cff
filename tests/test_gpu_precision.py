"""The precision form (operator contractions in complex FP32 on MFMA, QR,
norms and overlap terms in FP64) and its FP64 refinement of the last RK4
stage (mps_ctx_set_refine_sweeps), on d = 9 Rydberg-ladder density matrices
at D = 64, whose sites (36864 elements) take the three-stream path.

* The refinement's semantics: with FP64 contractions, a device RK4 step with
  n refinement half-sweeps equals the oracle's RK4 step followed by n calls
  of compress_sum_var_apply_H_left! / _right! with stage 4's arguments
  (compression.jl:889-944, 1007-1063), site by site to 1e-10, from a
  full-rank random state (no rank-deficient QR).
* The precision bar: 25 steps of FP32 + 2 refinement half-sweeps from the
  ground state against the FP64 oracle's plain RK4: per-site and total
  excited-level populations (and <sigma_z> = <n_e> - <n_g> per site) within
  1e-6, 1 - fidelity within 1e-12.  Both sides use the QR rank cutoff 1e-13
  (the ground state is zero-padded to D).
* Consecutive 11-argument calls in FP32 with writing entry points in between
  (the environments a call builds for the next one must then be dropped),
  with and without refinement, against the oracle at FP32 tolerance.
"""
import numpy as np
import pytest

import mpsamd as M
from mpsamd import models
from oracle import mps_oracle as O
from helpers import copy_sites, rel

pytestmark = pytest.mark.gpu
DT = 0.005
C4 = [-1.0 / 3, 1.0 / 3, 2.0 / 3, 1.0 / 3]


def _model(N, seed, n_exp=4):
    p = models.RydbergParams(n_exp=n_exp)
    z = models.positions(N, float(N), seed)
    return p, models.rydberg_rk4_mpos(p, z, DT), [1] + [p.w] * (N - 1) + [1]


def _oracle_step(L, rho, r, e, eop, nref=0):
    """One oracle RK4 step (11-argument form into a fresh state) and nref
    refit half-sweeps of stage 4; returns the new state."""
    out = copy_sites(rho)
    O.RK4stp_apply_H_half_(DT, L, rho, *r, *e, eop, psi0=out)
    A4 = [rho, *r]
    for k in range(nref):
        if k % 2 == 0:
            tmp = copy_sites(out)
            O.compress_sum_var_apply_H_left_(tmp, e, eop, out, C4, A4, L[2])
        else:
            O.compress_sum_var_apply_H_right_(out, e, eop, tmp, C4, A4, L[2])
    return out


def _device(rho, dims, wop, L, ctx):
    g = M.MPS.from_host(rho, ctx)
    gr = [M.MPS.from_host(rho, ctx) for _ in range(3)]
    ge = [M.build_env(dims, dims, ctx=ctx) for _ in range(3)]
    geop = M.build_env(dims, dims, wop, ctx=ctx)
    return g, gr, ge, geop, [M.MPO(h, ctx) for h in L]


@pytest.mark.parametrize("eleven", [False, True])
def test_rk4_refine_matches_oracle_refit(eleven):
    N, D, steps, nref = 6, 64, 2, 2
    print("case refine semantics eleven", eleven, flush=True)
    p, L, wop = _model(N, 7)
    rho, dims = O.mpsrandom(N, D, 9, np.random.default_rng(31))
    ref = copy_sites(rho)
    r = [copy_sites(rho) for _ in range(3)]
    e = [O.build_env(dims, dims) for _ in range(3)]
    eop = O.build_env_op(dims, dims, wop)
    for _ in range(steps):
        ref = _oracle_step(L, ref, r, e, eop, nref)
    ctx = M.Context(0)
    try:
        ctx.set_refine_sweeps(nref)
        g, gr, ge, geop, gL = _device(rho, dims, wop, L, ctx)
        go = M.MPS.from_host(rho, ctx) if eleven else None
        for _ in range(steps):
            if eleven:
                M.RK4stp_apply_H_half_(DT, gL, go, g, *gr, *ge, geop)
                g.copy_from(go)
            else:
                M.RK4stp_apply_H_half_(DT, gL, g, *gr, *ge, geop)
        got = g.to_host()
        # without refinement the device differs from the refitted oracle state
        ctx.set_refine_sweeps(0)
        h, hr, he, heop, hL = _device(rho, dims, wop, L, ctx)
        for _ in range(steps):
            M.RK4stp_apply_H_half_(DT, hL, h, *hr, *he, heop)
        plain = h.to_host()
    finally:
        ctx.synchronize()
    for j in range(N):
        assert rel(got[j], ref[j]) < 1e-10, j
    assert max(rel(plain[j], ref[j]) for j in range(N)) > 1e-12   # the refit changed the state


def _pops(sites_or_dev, N, ops, dev_ctx=None):
    """[tr, sum_j <op>] for each op, and per-site <op_j>, from scal_prod_no_conj."""
    if dev_ctx is not None:
        f = lambda m: M.scal_prod_no_conj(M.MPS.from_host(m, dev_ctx), sites_or_dev)   # noqa: E731
    else:
        f = lambda m: O.scal_prod_no_conj(m, sites_or_dev)   # noqa: E731
    tr = f(models.identity_vec_mps(N))
    tot = [f(models.sum_local_vec_mps(N, op)) / tr for op in ops]
    site = [[f(models.local_population_vec_mps(N, j, op)) / tr for j in range(N)] for op in ops]
    return tr, np.array(tot), np.array(site)


def test_rk4_fp32_refine_25_steps_against_fp64_oracle():
    N, D, steps, tol = 8, 64, 25, 1e-13
    p, L, wop = _model(N, 3)
    O.set_qr_rank_tol(tol)
    rho, dims = O.mpsgroundstate(N, D, 9)
    ref = copy_sites(rho)
    r = [copy_sites(rho) for _ in range(3)]
    e = [O.build_env(dims, dims) for _ in range(3)]
    eop = O.build_env_op(dims, dims, wop)
    print("case fp32+refine 25 steps: oracle", flush=True)
    for _ in range(steps):
        O.RK4stp_apply_H_half_(DT, L, ref, *r, *e, eop)
    ops = [models.HEE3, models.HSS3, models.HEE3 - models.HGG3]
    tr_o, tot_o, site_o = _pops(ref, N, ops)
    res = {}
    for nref in (0, 2):
        print("case fp32+refine 25 steps: device, refine", nref, flush=True)
        ctx = M.Context(0)
        ctx.set_qr_rank_tol(tol)
        ctx.set_precision("fp32")
        ctx.set_refine_sweeps(nref)
        g, gr, ge, geop, gL = _device(rho, dims, wop, L, ctx)
        for _ in range(steps):
            M.RK4stp_apply_H_half_(DT, gL, g, *gr, *ge, geop)
        ctx.set_precision("fp64")
        tr_d, tot_d, site_d = _pops(g, N, ops, ctx)
        ro = M.MPS.from_host(ref, ctx)
        fid = abs(M.scal_prod(g, ro)) ** 2 / (abs(M.scal_prod(g, g)) * abs(M.scal_prod(ro, ro)))
        res[nref] = {"trace": abs(tr_d - tr_o), "n_e": abs(tot_d[0] - tot_o[0]), "n_r": abs(tot_d[1] - tot_o[1]),
                     "site": float(np.abs(site_d - site_o).max()), "infid": 1.0 - fid}
        print("refine", nref, res[nref], flush=True)
    assert tot_o[0].real > 1e-4                       # the chain got excited
    r2 = res[2]
    assert r2["n_e"] < 1e-6 and r2["site"] < 1e-6 and r2["infid"] < 1e-12
    assert res[0]["n_e"] > 1e-13                      # the FP32 contractions did run


@pytest.mark.parametrize("nref", [0, 2])
def test_rk4_fp32_eleven_arg_consecutive_with_writes(nref):
    N, D, steps = 6, 64, 3
    print("case fp32 11-arg consecutive, refine", nref, flush=True)
    p, L, wop = _model(N, 5)
    rho, dims = O.mpsrandom(N, D, 9, np.random.default_rng(41))
    op = np.diag([1.0, 1.0 + 1e-3, 1.0]).astype(np.complex128)
    op9 = np.kron(op, np.eye(3))
    ref = copy_sites(rho)
    r = [copy_sites(rho) for _ in range(3)]
    e = [O.build_env(dims, dims) for _ in range(3)]
    eop = O.build_env_op(dims, dims, wop)
    for k in range(steps):
        ref = _oracle_step(L, ref, r, e, eop, nref)
        if k == 0:
            O.apply_site_operator_(ref[2], op9)
    ctx = M.Context(0)
    ctx.set_precision("fp32")
    ctx.set_refine_sweeps(nref)
    g, gr, ge, geop, gL = _device(rho, dims, wop, L, ctx)
    go = M.MPS.from_host(rho, ctx)
    for k in range(steps):
        M.RK4stp_apply_H_half_(DT, gL, go, g, *gr, *ge, geop)
        M.scal_prod(go, go)                 # read only
        g.copy_from(go)                     # a write: the next call's prepared environments are stale
        if k == 0:
            M.apply_site_operator_(g, 2, op9)
    ctx.set_precision("fp64")
    ro = M.MPS.from_host(ref, ctx)
    gg, rr = abs(M.scal_prod(g, g)), abs(M.scal_prod(ro, ro))
    infid = 1.0 - abs(M.scal_prod(g, ro)) ** 2 / (gg * rr)
    print("1 - fidelity", infid, "norm2", gg, rr, flush=True)
    assert infid < 1e-9 and abs(gg - rr) < 1e-5 * rr
