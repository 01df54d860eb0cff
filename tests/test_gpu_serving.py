"""Batched trajectories and the continuously batched server: a request's result
is the same alone and in the mix, and matches the FP64 oracle."""
import numpy as np
import pytest

from mpsamd import models
from mpsamd import serving as S
from oracle import trajectory as OT

pytestmark = pytest.mark.gpu

SAME = 1e-12   # alone vs in the mix (separate streams, fixed-order reductions: expected bit-identical)


def _req(rid, kind, N, D, steps, seed, arrival=0.0):
    return S.Request(rid, kind, N, D, steps, seed, arrival)


def test_batch_result_same_alone_and_in_the_mix():
    alone = [_req(0, "spin", 12, 8, 12, 7)]
    S.run_batch(alone)
    mix = [_req(0, "spin", 12, 8, 12, 7), _req(1, "spin", 16, 16, 9, 3), _req(2, "rydberg", 6, 9, 5, 2),
           _req(3, "spin", 10, 4, 15, 5)]
    S.run_batch(mix)
    d = np.abs(mix[0].result - alone[0].result).max()
    print("alone vs mix max difference", d, flush=True)
    assert d <= SAME
    seq = [_req(i, r.kind, r.N, r.D, r.steps, r.seed) for i, r in enumerate(mix)]
    S.run_batch(seq, sequential=True)
    for a, b in zip(seq, mix):
        assert np.abs(a.result - b.result).max() <= SAME


def test_spin_request_matches_oracle():
    # full bond dimension (N = 8, D = 16): the device trajectory, the oracle
    # trajectory and their sigma_z profiles agree within 1e-10
    r = _req(0, "spin", 8, 16, 20, 11)
    S.run_batch([r, _req(1, "rydberg", 5, 9, 4, 1)])
    from oracle import mps_oracle as O
    O.set_qr_rank_tol(S.QR_RANK_TOL)       # the server's per-request contexts use the cutoff
    z = models.positions(8, 8.0, 11)
    H = models.spin_rk4_mpos(z, 0.0, 0.01)
    jr, jl = models.spin_jump_mpos(z)
    ref = OT.OracleSpinTrajectory(8, 16, 11, H, jr, jl, 0.01)
    for _ in range(20):
        ref.step()
    assert np.abs(r.result[:-1] - ref.sigma_z()).max() < 1e-10
    assert abs(r.result[-1] - ref.p_coh[-1]) < 1e-10


def test_server_poisson_stream():
    reqs = S.poisson_requests(8, rate=50.0, seed=4, N_range=(8, 16), D_choices=(4, 8), steps_range=(4, 10),
                              rydberg_N_range=(4, 8), rydberg_D_choices=(9,))
    stats = S.Server(hbm_budget_bytes=2 << 30, max_active=4).run(reqs)
    print(stats, flush=True)
    assert stats["requests"] == len(reqs)
    assert 0 < stats["p50_latency_s"] <= stats["p99_latency_s"]
    # every response equals the same request solved alone
    for r in reqs:
        a = S.Request(r.rid, r.kind, r.N, r.D, r.steps, r.seed)
        S.run_batch([a])
        assert np.abs(a.result - r.result).max() <= SAME


# ------------------------------------------------------ lockstep trajectory pools
SYS = dict(dt=0.05, gam_eg=0.5, init=(0.0, 1.0))   # decay from |e...e>: jumps within the first steps


def _pool_run(pool, schedule, ticks):
    """schedule: {tick: [(request, slot)]}; runs ``ticks`` ticks."""
    done = []
    for t in range(ticks):
        for r, b in schedule.get(t, []):
            pool.admit(r, b)
        if pool.busy():
            pool.launch()
            done += pool.finish()
    return done


def test_pool_result_same_alone_and_in_the_mix():
    N, D, ps = 10, 8, 3
    alone = S.TrajRequest(0, N, D, ps, seed=77, steps=12)
    pool = S.TrajectoryPool(N, D, slots=6, **SYS)
    _pool_run(pool, {0: [(alone, 0)]}, 12)
    pool.close()
    # the same request in slot 4, admitted at tick 3 next to others that come
    # and go, some on other chains of the same shape
    rs = [S.TrajRequest(i, N, D, ps if i % 2 else 20 + i, seed=100 + i, steps=5 + 2 * i) for i in range(5)]
    mix = S.TrajRequest(9, N, D, ps, seed=77, steps=12)
    pool = S.TrajectoryPool(N, D, slots=6, **SYS)
    done = _pool_run(pool, {0: [(rs[0], 0), (rs[1], 1)], 2: [(rs[2], 2)], 3: [(mix, 4), (rs[3], 3)],
                            6: [(rs[4], 0)]}, 20)
    pool.close()
    assert mix in done and alone.result is not None
    d = np.abs(mix.result - alone.result).max()
    print("pool alone vs mix max difference", d, flush=True)
    assert d <= SAME
    # and the lockstep copy agrees with the trajectory on a plain context
    single = S.solve_alone(S.TrajRequest(0, N, D, ps, seed=77, steps=12), **SYS)
    assert np.abs(single.result - alone.result).max() < 1e-10


def test_pool_request_matches_oracle():
    # full bond dimension (N = 6, D = 8): every compression exact
    N, D, ps, seed, steps = 6, 8, 5, 31, 15
    r = S.TrajRequest(0, N, D, ps, seed=seed, steps=steps)
    pool = S.TrajectoryPool(N, D, slots=3, **SYS)
    _pool_run(pool, {0: [(S.TrajRequest(1, N, D, 9, seed=1, steps=4), 0)], 1: [(r, 2)]}, steps + 1)
    pool.close()
    from oracle import mps_oracle as O
    O.set_qr_rank_tol(S.QR_RANK_TOL)
    z = models.positions(N, float(N), ps)
    H = models.spin_rk4_mpos(z, 0.0, SYS["dt"])
    jr, jl = models.spin_jump_mpos(z)
    ref = OT.OracleSpinTrajectory(N, D, seed, H, jr, jl, SYS["dt"], SYS["gam_eg"], init=SYS["init"])
    for _ in range(steps):
        ref.step()
    assert ref.jumps, "the case is meant to take jumps"
    assert np.abs(r.result[:-1] - ref.sigma_z()).max() < 1e-10
    assert abs(r.result[-1] - ref.p_coh[-1]) < 1e-10


def test_pooled_server_poisson_stream():
    # three chains, two shapes: (8, 8) chains 1 and 5 share a pool
    systems = [(8, 8, 1), (10, 8, 2), (8, 8, 5)]
    reqs = S.poisson_trajectory_requests(16, rate=200.0, seed=6, systems=systems, steps_range=(3, 9))
    stats = S.PooledServer(slots=4, max_pools=2, **SYS).run(reqs, warm=[(8, 8)])
    print(stats, flush=True)
    assert stats["requests"] == len(reqs) and stats["pools"] == 2
    assert 0 < stats["p50_latency_s"] <= stats["p99_latency_s"]
    assert len({r.pos_seed for r in reqs if (r.N, r.D) == (8, 8)}) == 2
    # every response equals the same request run alone in slot 0 of a fresh pool
    for r in reqs:
        a = S.TrajRequest(r.rid, r.N, r.D, r.pos_seed, r.seed, r.steps)
        pool = S.TrajectoryPool(r.N, r.D, slots=4, **SYS)
        _pool_run(pool, {0: [(a, 0)]}, r.steps)
        pool.close()
        assert np.abs(a.result - r.result).max() <= SAME, r.rid


# ------------------------------------------------ lockstep density-matrix pools
def test_density_matrix_pool_same_alone_and_in_the_mix():
    N, D, ps, steps = 6, 9, 3, 6
    alone = S.TrajRequest(0, N, D, ps, seed=0, steps=steps, kind="rydberg")
    pool = S.DensityMatrixPool(N, D, slots=4)
    _pool_run(pool, {0: [(alone, 0)]}, steps)
    pool.close()
    # the same request in slot 2, admitted at tick 2 next to requests on
    # other chains of the same shape that come and go
    rs = [S.TrajRequest(i, N, D, 10 + i, seed=0, steps=3 + i, kind="rydberg") for i in range(4)]
    mix = S.TrajRequest(9, N, D, ps, seed=0, steps=steps, kind="rydberg")
    pool = S.DensityMatrixPool(N, D, slots=4)
    done = _pool_run(pool, {0: [(rs[0], 0), (rs[1], 1)], 2: [(mix, 2), (rs[2], 3)], 4: [(rs[3], 0)]}, 12)
    pool.close()
    assert mix in done and all(r in done for r in rs)
    d = np.abs(mix.result - alone.result).max()
    print("density-matrix pool alone vs mix max difference", d, mix.result, flush=True)
    assert d <= SAME
    # and the lockstep copy agrees with the evolution on a plain context
    single = S.solve_alone(S.TrajRequest(0, N, D, ps, seed=0, steps=steps, kind="rydberg"))
    assert np.abs(single.result - alone.result).max() < 1e-10


def test_precision_form_pool_same_alone_and_in_the_mix():
    # density matrices in the precision form (FP32 operator contractions, two
    # FP64 refinement half-sweeps per step) served from a lockstep pool: the
    # same request alone and in a mix gives the same bits; against the same
    # form on a plain context (other GEMM plans) and against the FP64 path it
    # agrees to the precision form's bar, and it differs from FP64 (FP32 ran)
    N, D, ps, steps = 6, 9, 3, 6
    print("case precision-form pool", N, D, steps, flush=True)
    alone = S.TrajRequest(0, N, D, ps, seed=0, steps=steps, kind="rydberg_p")
    pool = S.DensityMatrixPool(N, D, slots=4, precision="fp32", refine=2)
    _pool_run(pool, {0: [(alone, 0)]}, steps)
    pool.close()
    rs = [S.TrajRequest(i, N, D, 10 + i, seed=0, steps=3 + i, kind="rydberg_p") for i in range(4)]
    mix = S.TrajRequest(9, N, D, ps, seed=0, steps=steps, kind="rydberg_p")
    pool = S.DensityMatrixPool(N, D, slots=4, precision="fp32", refine=2)
    done = _pool_run(pool, {0: [(rs[0], 0), (rs[1], 1)], 2: [(mix, 2), (rs[2], 3)], 4: [(rs[3], 0)]}, 12)
    pool.close()
    assert mix in done and all(r in done for r in rs)
    assert np.abs(mix.result - alone.result).max() <= SAME
    plain_p = S.solve_alone(S.TrajRequest(0, N, D, ps, seed=0, steps=steps, kind="rydberg_p"))
    fp64 = S.solve_alone(S.TrajRequest(0, N, D, ps, seed=0, steps=steps, kind="rydberg"))
    d_plain = np.abs(plain_p.result - alone.result).max()
    d64 = np.abs(fp64.result - alone.result)   # trace and Rydberg population: the form's bar is absolute
    print("precision pool vs plain precision form", d_plain, "vs FP64", d64, "FP64", fp64.result, flush=True)
    assert d_plain < 1e-6
    assert d64.max() < 1e-6 and d64.max() > 1e-14


def test_density_matrix_pool_matches_oracle():
    N, D, ps, steps, dt = 5, 9, 4, 4, 0.005
    r = S.TrajRequest(0, N, D, ps, seed=0, steps=steps, kind="rydberg")
    pool = S.DensityMatrixPool(N, D, slots=3, dt=dt)
    _pool_run(pool, {0: [(S.TrajRequest(1, N, D, 8, seed=0, steps=2, kind="rydberg"), 0)], 1: [(r, 1)]},
              steps + 1)
    got = pool.rv[1].to_host()          # slot 1 after the request's last step
    pool.close()
    from oracle import mps_oracle as O
    from oracle import dense as Dn
    from helpers import copy_sites, rel
    O.set_qr_rank_tol(S.QR_RANK_TOL)
    try:
        p = models.RydbergParams(n_exp=2)
        L = models.rydberg_rk4_mpos(p, models.positions(N, float(N), ps), dt)
        g = np.zeros(9, dtype=complex)
        g[0] = 1.0
        rho, dims = O.mpsproductstate(N, D, 9, g)
        st = [copy_sites(rho) for _ in range(3)]
        envs = [O.build_env(dims, dims) for _ in range(3)]
        eop = O.build_env_op(dims, dims, [1] + [p.w] * (N - 1) + [1])
        for _ in range(steps):
            O.RK4stp_apply_H_half_(dt, L, rho, *st, *envs, eop)
    finally:
        O.set_qr_rank_tol(0.0)
    tr = O.scal_prod_no_conj(models.identity_vec_mps(N), rho)
    ne = O.scal_prod_no_conj(models.sum_local_vec_mps(N, models.HSS3), rho) / tr
    print("pool", r.result, "oracle", tr.real, ne.real, flush=True)
    assert rel(Dn.mps_to_dense(got), Dn.mps_to_dense(rho)) < 1e-10
    assert abs(r.result[0] - tr.real) < 1e-10
    assert abs(r.result[1] - ne.real) < 1e-10 * max(1.0, abs(ne))


def test_pooled_server_mixes_trajectories_and_density_matrices():
    systems = [(8, 8, 1), ("rydberg", 5, 9, 2), ("rydberg", 5, 9, 6)]
    reqs = S.poisson_trajectory_requests(12, rate=200.0, seed=8, systems=systems, steps_range=(3, 7))
    kinds = {r.kind for r in reqs}
    assert kinds == {"spin", "rydberg"}
    stats = S.PooledServer(slots=4, max_pools=2, dm_slots=3, **SYS).run(reqs, warm=[("rydberg", 5, 9)])
    print(stats, flush=True)
    assert stats["requests"] == len(reqs) and stats["pools"] == 2
    # every response equals the same request run alone in slot 0 of a fresh pool
    for r in reqs:
        a = S.TrajRequest(r.rid, r.N, r.D, r.pos_seed, r.seed, r.steps, kind=r.kind)
        if r.kind == "rydberg":
            pool = S.DensityMatrixPool(r.N, r.D, slots=3)
        else:
            pool = S.TrajectoryPool(r.N, r.D, slots=4, **SYS)
        _pool_run(pool, {0: [(a, 0)]}, r.steps)
        pool.close()
        assert np.abs(a.result - r.result).max() <= SAME, (r.rid, r.kind)


def test_pooled_server_hbm_budget():
    # the pool sizes the server plans with are the slabs the pools allocate
    srv = S.PooledServer(slots=4, max_pools=4, dm_slots=3, **SYS)
    b1, b2 = srv.pool_bytes(("spin", 8, 8)), srv.pool_bytes(("spin", 10, 8))
    pool = S.TrajectoryPool(8, 8, 4, **SYS)
    assert pool.ens.hbm_bytes == b1
    pool.close()
    dm = S.DensityMatrixPool(5, 9, slots=3)
    assert dm.hbm_bytes == srv.pool_bytes(("rydberg", 5, 9))
    dm.close()
    # a budget that holds one of the two pools at a time: they take turns
    budget = max(b1, b2) + min(b1, b2) // 2
    systems = [(8, 8, 1), (10, 8, 2)]
    reqs = S.poisson_trajectory_requests(10, rate=200.0, seed=9, systems=systems, steps_range=(3, 6))
    assert {(r.N, r.D) for r in reqs} == {(8, 8), (10, 8)}
    stats = S.PooledServer(slots=4, max_pools=4, hbm_budget_bytes=budget, **SYS).run(reqs)
    print(stats, flush=True)
    assert stats["requests"] == len(reqs)
    assert stats["pools_opened"] >= 2
    assert stats["peak_pool_bytes"] <= budget
    for r in reqs:
        a = S.TrajRequest(r.rid, r.N, r.D, r.pos_seed, r.seed, r.steps)
        pool = S.TrajectoryPool(r.N, r.D, slots=4, **SYS)
        _pool_run(pool, {0: [(a, 0)]}, r.steps)
        pool.close()
        assert np.abs(a.result - r.result).max() <= SAME, r.rid
    # a pool larger than the whole budget is refused before anything runs
    with pytest.raises(ValueError):
        S.PooledServer(slots=4, hbm_budget_bytes=min(b1, b2) - 1, **SYS).run(reqs)
