"""Half-sweeps and RK4 at bond dimensions whose sites use the context's second
stream (sites of at least 32768 elements: the overlap terms run concurrently
with the operator chain).  Same FP64 oracle, same 1e-10 bar as the
single-stream sizes, and the same result as the single-stream schedule that
kernel timing forces (the sums are accumulated in the same order)."""
import numpy as np
import pytest

import mpsamd as M
from mpsamd import models
from oracle import mps_oracle as O
from helpers import rel, copy_sites, random_mpo, crandn

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.mark.parametrize("side", ["left", "right"])
@pytest.mark.parametrize("numv", [2, 4])
def test_compress_sum_var_apply_H_two_streams(side, numv, rng):
    N, D, d, w = 6, 64, 9, 4          # middle sites 64 x 9 x 64 = 36864 elements
    print("case", side, numv, flush=True)
    states = [O.mpsrandom(N, D, d, rng)[0] for _ in range(numv)]
    Ain, dims = O.mpsrandom(N, D, d, rng)
    if side == "right":
        O.rightorthonormalizationQR_(Ain)
    H = random_mpo(rng, N, d, w)
    coef = list(crandn(rng, numv))
    wd = [1] + [w] * (N - 1) + [1]
    A0 = copy_sites(Ain)
    env = [O.build_env(dims, dims) for _ in range(numv - 1)]
    envop = O.build_env_op(dims, dims, wd)
    fn = O.compress_sum_var_apply_H_left_ if side == "left" else O.compress_sum_var_apply_H_right_
    fn(A0, env, envop, Ain, coef, states, H)

    gfn = M.compress_sum_var_apply_H_left_ if side == "left" else M.compress_sum_var_apply_H_right_
    ctx = M.default_context()
    out = []
    for timed in (False, True):        # kernel timing runs everything on one stream
        ctx.set_profiling(timed)
        g_states = [M.MPS.from_host(s) for s in states]
        g_A0 = M.MPS.from_host(Ain)
        g_env = [M.build_env(dims, dims) for _ in range(numv - 1)]
        g_envop = M.build_env(dims, dims, wd)
        gfn(g_A0, g_env, g_envop, M.MPS.from_host(Ain), coef, g_states, M.MPO(H))
        out.append(g_A0.to_host())
        ctx.set_profiling(False)
    for j in range(N):
        assert rel(out[0][j], A0[j]) < TOL, j
        assert np.array_equal(out[0][j], out[1][j]), j


@pytest.mark.parametrize("N,eleven", [(6, False), (6, True), (12, False)])
def test_rk4_rydberg_two_streams(N, eleven):
    D, dt, steps = 64, 0.005, 2
    print("case", N, "eleven", eleven, flush=True)
    p = models.RydbergParams(n_exp=2)
    z = models.positions(N, float(N), 7)
    L = models.rydberg_rk4_mpos(p, z, dt)
    rho, dims = O.mpsrandom(N, D, 9, np.random.default_rng(11))
    wop = [1] + [p.w] * (N - 1) + [1]
    # oracle
    ref = copy_sites(rho)
    envop = O.build_env_op(dims, dims, wop)
    e = [O.build_env(dims, dims) for _ in range(3)]
    r1, r2, r3 = copy_sites(rho), copy_sites(rho), copy_sites(rho)
    for _ in range(steps):
        if eleven:
            o = copy_sites(ref)
            O.RK4stp_apply_H_half_(dt, L, ref, r1, r2, r3, *e, envop, psi0=o)
            ref = o
        else:
            O.RK4stp_apply_H_half_(dt, L, ref, r1, r2, r3, *e, envop)
    # device, two streams
    gL = [M.MPO(h) for h in L]
    g = M.MPS.from_host(rho)
    g1, g2, g3 = (M.MPS.from_host(rho) for _ in range(3))
    ge = [M.build_env(dims, dims) for _ in range(3)]
    geop = M.build_env(dims, dims, wop)
    go = M.MPS.from_host(rho) if eleven else None
    for _ in range(steps):
        if eleven:
            M.RK4stp_apply_H_half_(dt, gL, go, g, g1, g2, g3, *ge, geop)
            g.copy_from(go)
        else:
            M.RK4stp_apply_H_half_(dt, gL, g, g1, g2, g3, *ge, geop)
    got = g.to_host()
    for j in range(N):
        assert rel(got[j], ref[j]) < TOL, j
    I = models.identity_vec_mps(N)
    assert abs(M.scal_prod_no_conj(M.MPS.from_host(I), g) - O.scal_prod_no_conj(I, ref)) < TOL


def test_right_gauge_moves_large_sites(rng):
    # rightorth! on sites of at least 32768 elements factors a transposed,
    # column-major copy of the panel: every site against the oracle
    N, D, d = 5, 64, 9
    a, dims = O.mpsrandom(N, D, d, rng)
    a = [s + 0.1 * crandn(rng, s.shape) for s in a]
    ref = copy_sites(a)
    no_ref = O.rightorthonormalizationQR_(ref)
    g = M.MPS.from_host(a)
    assert abs(M.rightorthonormalizationQR_(g) - no_ref) < TOL * no_ref
    got = g.to_host()
    for j in range(N):
        assert rel(got[j], ref[j]) < TOL, j


def test_rk4_consecutive_calls_reuse_and_invalidation():
    # Each RK4 call at these sizes queues the next call's first environments
    # behind its last sweep; they are used only when nothing was written in
    # between.  Sequence: step, read-only query (kept), step, a write to the
    # state (normalize_lo!: dropped), step, a different envop (dropped), step.
    N, D, dt = 6, 64, 0.005
    p = models.RydbergParams(n_exp=2)
    z = models.positions(N, 6.0, 5)
    L = models.rydberg_rk4_mpos(p, z, dt)
    rho, dims = O.mpsrandom(N, D, 9, np.random.default_rng(21))
    wop = [1] + [p.w] * (N - 1) + [1]
    ref = copy_sites(rho)
    envop = O.build_env_op(dims, dims, wop)
    e = [O.build_env(dims, dims) for _ in range(3)]
    r1, r2, r3 = copy_sites(rho), copy_sites(rho), copy_sites(rho)
    gL = [M.MPO(h) for h in L]
    g = M.MPS.from_host(rho)
    g1, g2, g3 = (M.MPS.from_host(rho) for _ in range(3))
    ge = [M.build_env(dims, dims) for _ in range(3)]
    geop = M.build_env(dims, dims, wop)
    geop2 = M.build_env(dims, dims, wop)
    I = models.identity_vec_mps(N)
    gI = M.MPS.from_host(I)

    def both(envop_dev):
        O.RK4stp_apply_H_half_(dt, L, ref, r1, r2, r3, *e, envop)
        M.RK4stp_apply_H_half_(dt, gL, g, g1, g2, g3, *ge, envop_dev)
        got = g.to_host()
        for j in range(N):
            assert rel(got[j], ref[j]) < TOL, j

    both(geop)
    assert abs(M.scal_prod_no_conj(gI, g) - O.scal_prod_no_conj(I, ref)) < TOL
    both(geop)
    assert abs(M.normalize_lo_(g) - O.normalize_lo_(ref)) < TOL
    both(geop)
    # a write to site 1 changes every left environment past it
    op = np.eye(9, dtype=complex) + 0.3 * crandn(np.random.default_rng(3), (9, 9))
    ref[1] = O.apply_site_operator_(ref[1], op)
    M.apply_site_operator_(g, 1, op)
    both(geop)
    both(geop2)
    both(geop2)


def test_rk4_fp32_two_streams():
    # the precision form on the three-stream path: consecutive calls (the
    # second reuses environments built behind the first) stay within the FP32
    # level of the FP64 oracle and above FP64 round-off
    N, D, dt, steps = 6, 64, 0.005, 2
    p = models.RydbergParams(n_exp=2)
    z = models.positions(N, float(N), 7)
    L = models.rydberg_rk4_mpos(p, z, dt)
    rho, dims = O.mpsrandom(N, D, 9, np.random.default_rng(13))
    wop = [1] + [p.w] * (N - 1) + [1]
    ref = copy_sites(rho)
    envop = O.build_env_op(dims, dims, wop)
    e = [O.build_env(dims, dims) for _ in range(3)]
    r1, r2, r3 = copy_sites(rho), copy_sites(rho), copy_sites(rho)
    for _ in range(steps):
        O.RK4stp_apply_H_half_(dt, L, ref, r1, r2, r3, *e, envop)
    ctx = M.default_context()
    ctx.set_precision("fp32")
    try:
        gL = [M.MPO(h) for h in L]
        g = M.MPS.from_host(rho)
        g1, g2, g3 = (M.MPS.from_host(rho) for _ in range(3))
        ge = [M.build_env(dims, dims) for _ in range(3)]
        geop = M.build_env(dims, dims, wop)
        for _ in range(steps):
            M.RK4stp_apply_H_half_(dt, gL, g, g1, g2, g3, *ge, geop)
        got = g.to_host()
    finally:
        ctx.set_precision("fp64")
    err = max(rel(got[j], ref[j]) for j in range(N))
    print("fp32 two-stream relative site error", err, flush=True)
    assert 1e-12 < err < 1e-5


def test_arnoldi_two_streams(rng):
    # Arnoldi_expm! at D = 64, d = 9 (its variational sums on two streams)
    N, D, d, numv = 5, 64, 9, 3
    psi, dims = O.mpsrandom(N, D, d, rng)
    psi = [s.real.astype(complex) for s in psi]
    O.leftorthonormalizationQR_(psi)
    H = [np.real(h).astype(complex) * 0.3 for h in random_mpo(rng, N, d, 2)]
    wd = [1] + [2] * (N - 1) + [1]
    rhov = [copy_sites(psi) for _ in range(numv)]
    envop = O.build_env_op(dims, dims, wd)
    cref = O.Arnoldi_expm_(H, rhov, numv, None, envop)
    grhov = [M.MPS.from_host(psi) for _ in range(numv)]
    cgot = M.Arnoldi_expm_(M.MPO(H), grhov, numv, None, M.build_env(dims, dims, wd))
    assert np.abs(cgot - cref).max() < 1e-9
    got = grhov[0].to_host()
    for j in range(N):
        assert rel(got[j], rhov[0][j]) < 1e-9, j


def test_rk4_capped_environment_stream_bitwise(monkeypatch):
    # The stream that builds the next sweep's environments runs its GEMMs on
    # a capped grid (workgroups walk several output blocks); the per-block
    # arithmetic is unchanged, so a hard cap (8 workgroups) and no cap give
    # the same bits after two steps.
    N, D, dt = 6, 64, 0.005
    p = models.RydbergParams(n_exp=2)
    z = models.positions(N, 6.0, 3)
    L = models.rydberg_rk4_mpos(p, z, dt)
    rho, dims = O.mpsrandom(N, D, 9, np.random.default_rng(31))
    wop = [1] + [p.w] * (N - 1) + [1]
    out = []
    for cap in ("0", "8"):
        print("case aux cap", cap, flush=True)
        monkeypatch.setenv("MPS_AUX_MAX_WG", cap)   # read when the context first needs the stream
        ctx = M.Context(0)
        gL = [M.MPO(h, ctx) for h in L]
        g = M.MPS.from_host(rho, ctx)
        g1, g2, g3 = (M.MPS.from_host(rho, ctx) for _ in range(3))
        ge = [M.build_env(dims, dims, ctx=ctx) for _ in range(3)]
        geop = M.build_env(dims, dims, wop, ctx=ctx)
        for _ in range(2):
            M.RK4stp_apply_H_half_(dt, gL, g, g1, g2, g3, *ge, geop)
        out.append(g.to_host())
        for o in (g, g1, g2, g3, *ge, geop, *gL):
            o.__del__()
        ctx.close()
    for j in range(N):
        assert np.array_equal(out[0][j], out[1][j]), j
