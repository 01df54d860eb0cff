"""Benchmark: RK4 sweeps/s of the density-matrix workload on one MI355X.

Workload (the reference's GPU density-matrix loop, rydberg_density_matrix_gpu.jl:211-222):
three-level Rydberg ladder, vectorised Lindbladian (d = 9), N = 100 atoms at
seeded uniform positions in [0, 100], C6/r^6 fitted by 4 exponentials
(MPO bond w = 14), D_max = 128, dt = 0.005, starting from the ground-state
product MPS (mpsgroundstate).  One step = one RK4stp_apply_H_half! call
(10-argument form, result overwrites rho).

Prints one JSON line.  ``value`` is FP64 RK4 sweeps/s with all state
resident in HBM.  ``roofline`` comes from a second pass over the same steps
in which every GEMM launch is bracketed by HIP events on the library's
stream (dominant kernel: the complex FP64 MFMA GEMM).  ``cpu_baseline`` is
the numpy/OpenBLAS oracle timed on this host over one full RK4 step of the
same workload.
"""
import argparse
import json
import os
import sys
import time

# more hardware queues than HIP's default of 4 for the concurrent solves of the
# batched / serving legs (read once, at the first HIP call of the process)
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "matrixproductstates.jl_amd"))

import numpy as np  # noqa: E402

METRIC = "RK4 sweeps/s at N=100,D=128 (FP64 vs FP32+refine); p50/p99 trajectory latency"
PEAK = {"fp64": 78.6, "fp32": 157.3}   # TFLOP/s, MI355X dense MFMA


def model_flops(dims, d, w):
    """Algorithmic GEMM flops of one RK4 half step (8 flops / complex MAC), from the
    reference's contraction shapes: per site 20 w d Dl Dr D + 8 w^2 d^2 Dl Dr + 25 d D^3-type terms."""
    tot = 0.0
    N = len(dims) - 1
    for j in range(N):
        Dl, Dr = dims[j], dims[j + 1]
        wl = 1 if j == 0 else w
        wr = 1 if j == N - 1 else w
        wm = max(wl, wr)
        # 5 op-term GEMMs of (Dl d) x (wm D) x D per stage x 4 stages  ~ 20 w d D^3
        op = 20.0 * wm * d * Dl * Dr * max(Dl, Dr)
        mpo = 8.0 * wl * wr * d * d * Dl * Dr
        ovl = 25.0 * d * Dl * Dr * max(Dl, Dr)
        tot += op + mpo + ovl
    return 8.0 * tot


QR_RANK_TOL = 1e-13   # the cutoff the full-size parity tests use on both sides

# Counter measurements of the dominant GEMM launch at this workload's shape, the
# 1792x1152x128 product of prod_AOL (complex FP64, 64-tiles), from rocprofv3
# over `bench.py --gemm-counters fp64` (FETCH_SIZE x 2.00 and WRITE_SIZE x 1.00,
# the factors a 256 MiB copy gives; SQ_VALU_MFMA_BUSY_CYCLES over the kernel's
# cycles at 2.4 GHz): profiles/zgemm_prodAOL_N100_D128_d9_w14_counters.md
DOMINANT_LAUNCH = {"kernel": "zgemm_f64_kernel<64>", "shape": "1792x1152x128 (prod_AOL, D=128, d=9, w=14)",
                   "hbm_read_bytes": 26.55e6, "hbm_write_bytes": 33.03e6, "algorithmic_bytes": 39.06e6,
                   "kernel_us": 42.4, "mfma_busy": 0.48}


def setup(N, D, n_exp, dt, seed, ctx):
    import mpsamd as M
    from mpsamd import models
    ctx.set_qr_rank_tol(QR_RANK_TOL)
    p = models.RydbergParams(n_exp=n_exp)
    z = models.positions(N, float(N), seed)
    L = [M.MPO(h, ctx) for h in models.rydberg_rk4_mpos(p, z, dt)]
    rho, dims = M.mpsgroundstate(N, D, 9, ctx)
    rhos = [rho.similar() for _ in range(3)]
    for r in rhos:
        r.copy_from(rho)
    envs = [M.build_env(dims, dims, ctx=ctx) for _ in range(3)]
    envop = M.build_env(dims, dims, [1] + [p.w] * (N - 1) + [1], ctx=ctx)
    return p, z, L, rho, rhos, envs, envop, dims


def cpu_baseline(N, D, n_exp, dt, seed, steps_cpu=1):
    """The numpy/OpenBLAS oracle on this host: one full RK4 step of the same
    workload at every BLAS thread (the reported value), then stage 1 alone
    (compress_var_apply_H_left!) at 2 BLAS threads and at every thread, the
    reference's own BLAS.set_num_threads(2) setting
    (examples/rydberg_density_matrix.jl:3), kept to a bounded sample."""
    from oracle import mps_oracle as O
    from mpsamd import models
    try:
        from threadpoolctl import threadpool_info, threadpool_limits
        threads = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        threadpool_limits = None
        threads = os.cpu_count()
    try:
        import psutil
        phys = psutil.cpu_count(logical=False)
    except Exception:
        phys = None
    p = models.RydbergParams(n_exp=n_exp)
    z = models.positions(N, float(N), seed)
    L = models.rydberg_rk4_mpos(p, z, dt)
    O.set_qr_rank_tol(QR_RANK_TOL)
    rho, dims = O.mpsgroundstate(N, D, 9)
    wop = [1] + [p.w] * (N - 1) + [1]
    envop = O.build_env_op(dims, dims, wop)
    e = [O.build_env(dims, dims) for _ in range(3)]
    r1, r2, r3 = ([a.copy() for a in rho] for _ in range(3))
    t0 = time.perf_counter()
    for _ in range(steps_cpu):
        O.RK4stp_apply_H_half_(dt, L, rho, r1, r2, r3, *e, envop)
    t = time.perf_counter() - t0
    stage1 = {}
    for nt in ([2, threads] if threadpool_limits else []):
        with threadpool_limits(limits=nt, user_api="blas"):
            t1 = time.perf_counter()
            O.compress_var_apply_H_left_(r1, envop, rho, rho, L[0])
            stage1[str(nt)] = round(time.perf_counter() - t1, 2)
    O.set_qr_rank_tol(0.0)
    out = {"value": steps_cpu / t, "unit": "sweeps/s", "cores": int(threads), "kind": "port",
           "sample": f"{steps_cpu} full RK4stp_apply_H_half! step(s) of the same N={N} workload "
                     f"(numpy/OpenBLAS zgemm + LAPACK zgeqrf/zungqr) at {threads} BLAS threads, {t:.1f} s",
           "host_logical_cpus": os.cpu_count(), "host_physical_cores": phys}
    if stage1:
        out["stage1_s_by_blas_threads"] = stage1
        out["est_sweeps_per_s_2_threads"] = round(out["value"] * stage1[str(threads)] / stage1["2"], 5)
    print(f"cpu baseline: {out}", file=sys.stderr, flush=True)
    return out


def ensemble_leg(seed, N=60, D=32, B=256, steps=10, n_seq=4):
    """Lockstep ensemble of B quantum-jump trajectories of one system (one
    batched context: every kernel launched once for all B) against the same
    trajectories solved one after another on a plain context."""
    import mpsamd as M
    from mpsamd.trajectory import SpinEnsemble, SpinTrajectory
    ctx = M.Context(0)
    trs = [SpinTrajectory(N, D, seed + s, dt=0.01, ctx=ctx, pos_seed=seed) for s in range(n_seq)]
    for tr in trs:          # first two steps: eager, then graph capture
        tr.step(), tr.step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for tr in trs:
        for _ in range(steps):
            tr.step()
    ctx.synchronize()
    seq = (time.perf_counter() - t0) / (n_seq * steps)
    for tr in trs:
        for v in list(vars(tr).values()):
            for o in (v if isinstance(v, (list, tuple)) else [v]):
                if isinstance(o, (M.MPS, M.MPO, M.Env)):
                    o.__del__()
    ctx.close()
    ens = SpinEnsemble(N, D, seed, [seed + s for s in range(B)], dt=0.01)
    ens.step(), ens.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        ens.step()
    per = (time.perf_counter() - t0) / (B * steps)
    jumps = sum(len(j) for j in ens.jumps)
    ens.close()
    out = {"trajectories": B, "N": N, "D": D, "steps_timed": steps, "jumps": jumps,
           "trajectory_steps_per_s": round(1.0 / per, 1), "sequential_trajectory_steps_per_s": round(1.0 / seq, 1),
           "speedup_vs_sequential": round(seq / per, 2)}
    print(f"lockstep ensemble: {out}", file=sys.stderr, flush=True)
    return out


def density_matrix_pool_leg(seed, N=40, D=32, B=64, steps=10, n_seq=4):
    """Lockstep pool of B Rydberg density-matrix evolutions (d=9; one batched
    context, each copy on its own chain's positions) against the same
    evolutions one after another on a plain context."""
    import mpsamd as M
    from mpsamd import serving as S
    ctx = M.Context(0)
    ctx.set_qr_rank_tol(S.QR_RANK_TOL)
    evs = [S.RydbergEvolution(N, D, seed + s, ctx=ctx) for s in range(n_seq)]
    for ev in evs:          # first two steps: eager, then graph capture
        for _ in range(2):
            ev.launch_coherent()
            ev.finish_step()
    t0 = time.perf_counter()
    for ev in evs:
        for _ in range(steps):
            ev.launch_coherent()
            ev.finish_step()
    seq = (time.perf_counter() - t0) / (n_seq * steps)
    for ev in evs:
        for v in list(vars(ev).values()):
            for o in (v if isinstance(v, (list, tuple)) else [v]):
                if isinstance(o, (M.MPS, M.MPO, M.Env)):
                    o.__del__()
    ctx.close()
    pool = S.DensityMatrixPool(N, D, slots=B, pos_seed=seed)
    for b in range(B):
        pool.admit(S.TrajRequest(b, N, D, seed + b, 0, steps + 3, kind="rydberg"), b)   # none retires inside the clock
    for _ in range(2):
        pool.launch()
        pool.finish()
    t0 = time.perf_counter()
    for _ in range(steps):
        pool.launch()
        pool.finish()
    per = (time.perf_counter() - t0) / (B * steps)
    pool.close()
    out = {"evolutions": B, "N": N, "D": D, "d": 9, "steps_timed": steps,
           "evolution_steps_per_s": round(1.0 / per, 1), "sequential_evolution_steps_per_s": round(1.0 / seq, 1),
           "speedup_vs_sequential": round(seq / per, 2)}
    print(f"density-matrix pool: {out}", file=sys.stderr, flush=True)
    return out


def uneven_pools_leg(seed, n=512, steps=10, Ns=(40, 60, 80, 100, 120), Ds=(16, 32, 64), slots=64):
    """n quantum-jump trajectories of uneven shape (N and D drawn per
    trajectory, one chain per shape) as one burst through the pooled server:
    a lockstep pool per shape, the pools advancing concurrently on their own
    streams.  The sequential rate is measured on one trajectory per shape on a
    plain context and weighted by the number of trajectories of that shape."""
    import mpsamd as M
    from mpsamd import serving as S
    from mpsamd.trajectory import SpinTrajectory
    rng = np.random.default_rng(seed)
    shapes = [(N, D) for N in Ns for D in Ds]
    pick = rng.integers(len(shapes), size=n)
    reqs = [S.TrajRequest(i, *shapes[k], seed + k, int(rng.integers(1 << 30)), steps) for i, k in enumerate(pick)]
    st = S.PooledServer(slots=slots, max_pools=len(shapes), dt=0.01).run(reqs)
    print(f"uneven pools: {st}", file=sys.stderr, flush=True)
    ctx = M.Context(0)
    ctx.set_qr_rank_tol(S.QR_RANK_TOL)
    t_seq = 0.0
    for k, (N, D) in enumerate(shapes):
        tr = SpinTrajectory(N, D, seed, dt=0.01, ctx=ctx, pos_seed=seed + k)
        tr.step(), tr.step()       # eager, then graph capture
        t0 = time.perf_counter()
        for _ in range(steps):
            tr.step()
        t_seq += (time.perf_counter() - t0) * int((pick == k).sum())
        for v in list(vars(tr).values()):
            for o in (v if isinstance(v, (list, tuple)) else [v]):
                if isinstance(o, (M.MPS, M.MPO, M.Env)):
                    o.__del__()
    ctx.close()
    out = {"trajectories": n, "steps_each": steps, "N_choices": list(Ns), "D_choices": list(Ds),
           "pools": st["pools"], "slots_per_pool": slots, "wall_s": round(st["wall_s"], 3),
           "trajectories_per_s": round(n / st["wall_s"], 2),
           "sequential_trajectories_per_s": round(n / t_seq, 3),
           "speedup_vs_sequential": round(t_seq / st["wall_s"], 2),
           "note": "wall time includes building each pool and capturing its step graph"}
    print(f"uneven pools: {out}", file=sys.stderr, flush=True)
    return out


def gemm_counters(prec, D=128, d=9, w=14, reps=5):
    """For counter runs of the dominant GEMMs (run under rocprofv3): a 256 MiB
    device copy (calibration of the byte counters), then `reps` calls of
    prod_AOL (mps_base.jl:26-39) at the benchmark shape through the C-ABI:
    its 1792x1152x128 product and its batched MPO product.  torch provides
    the device buffers; the library computes."""
    import ctypes as C
    import torch
    torch.cuda.init()
    n = 256 << 20
    src = torch.ones(n // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    import mpsamd as M
    from mpsamd._lib import call
    ctx = M.default_context()
    ctx.set_precision(prec)
    rng = np.random.default_rng(0)

    def dev(*shape):
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        t = torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to("cuda")
        torch.cuda.synchronize()
        return t

    A, O, L = dev(D, d, D), dev(w, d, w, d), dev(D, w, D)
    out = torch.zeros(D * d * D * w, dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    p = lambda t: C.c_void_p(t.data_ptr())   # noqa: E731
    for _ in range(reps):
        call("mps_dev_prod_AOL", ctx.h, p(out), p(A), D, d, D, p(O), w, w, p(L), D)
    ctx.synchronize()
    return {"calibration_copy_bytes": n, "prod_AOL_calls": reps, "shape": {"D": D, "d": d, "w": w},
            "precision": prec}


def ragged_pools_leg(seed, n=512, steps=5, Ds=(16, 32, 64, 128, 256), N_range=(40, 120),
                     budget=240 << 30, pool_cap=96 << 30, precision="fp32"):
    """n quantum-jump trajectories of the driven spin chain with N uniform in
    N_range and D drawn from Ds: one ragged pool per D (chains of every
    length advanced by one RK4 call, contractions in complex FP32), split into
    pools of at most pool_cap bytes; pools run concurrently on their own
    streams while their slabs fit the HBM budget, the rest wait for one to
    finish.  The sequential rate is one trajectory after another on a plain
    context (N = 40 and 120 timed per D, linear in N in between), in FP64 (the
    plain path) and in FP32."""
    import mpsamd as M
    from mpsamd.core import _dims_rule
    from mpsamd.trajectory import RaggedSpinEnsemble, SpinTrajectory, ragged_ensemble_bytes
    rng = np.random.default_rng(seed)
    Ns = rng.integers(N_range[0], N_range[1] + 1, size=n)
    Dn = rng.choice(np.asarray(Ds), size=n)
    seeds = rng.integers(1 << 30, size=n)
    plans = []   # (D, trajectory indices), longest chains first within a D
    for D in sorted(Ds, reverse=True):
        idx = sorted(np.flatnonzero(Dn == D).tolist(), key=lambda i: -Ns[i])
        while idx:
            k = len(idx)
            while k > 2 and ragged_ensemble_bytes([Ns[i] for i in idx[:k]], D) > pool_cap:
                k = (k + 1) // 2
            plans.append((int(D), idx[:k]))
            idx = idx[k:]
    flops = {}
    for i in range(n):
        key = (int(Ns[i]), int(Dn[i]))
        if key not in flops:
            flops[key] = model_flops(_dims_rule(key[0], key[1], 2), 2, 4)
    total_flops = steps * sum(flops[(int(Ns[i]), int(Dn[i]))] for i in range(n))

    t0 = time.perf_counter()
    pending, active, used, peak, done = list(plans), [], 0, 0, 0
    jumps = 0
    while pending or active:
        while pending:
            D, idx = pending[0]
            need = ragged_ensemble_bytes([Ns[i] for i in idx], D)
            if active and used + need > budget:
                break
            pending.pop(0)
            ens = RaggedSpinEnsemble([int(Ns[i]) for i in idx], D, [seed + int(Ns[i]) for i in idx],
                                     [int(seeds[i]) for i in idx], dt=0.01, qr_rank_tol=1e-13,
                                     precision=precision)
            active.append([ens, 0, need])
            used += need
            peak = max(peak, used)
        for a in active:
            a[0].launch()
        for a in active:
            a[0].finish()
            a[1] += 1
        for a in [a for a in active if a[1] >= steps]:
            jumps += sum(len(j) for j in a[0].jumps)
            done += a[0].B
            a[0].close()
            used -= a[2]
            active.remove(a)
    wall = time.perf_counter() - t0
    print(f"ragged pools: {n} trajectories in {len(plans)} pools, {wall:.2f} s", file=sys.stderr, flush=True)

    def seq_rate(prec):
        ctx = M.Context(0)
        ctx.set_qr_rank_tol(1e-13)
        ctx.set_precision(prec)
        t_seq = 0.0
        for D in Ds:
            per = {}
            for N in N_range:
                tr = SpinTrajectory(int(N), int(D), seed, dt=0.01, ctx=ctx, pos_seed=seed + int(N))
                tr.step(), tr.step()       # eager, then graph capture
                ctx.synchronize()
                ts = time.perf_counter()
                for _ in range(2):
                    tr.step()
                per[N] = (time.perf_counter() - ts) / 2
                for v in list(vars(tr).values()):
                    for o in (v if isinstance(v, (list, tuple)) else [v]):
                        if isinstance(o, (M.MPS, M.MPO, M.Env)):
                            o.__del__()
            a, b = per[N_range[0]], per[N_range[1]]
            for i in np.flatnonzero(Dn == D):
                t_seq += steps * (a + (b - a) * (Ns[i] - N_range[0]) / (N_range[1] - N_range[0]))
        ctx.close()
        return t_seq

    t64, t32 = seq_rate("fp64"), seq_rate("fp32")
    out = {"trajectories": n, "steps_each": steps, "N_range": list(N_range), "D_choices": list(Ds),
           "precision": precision, "pools": len(plans),
           "pools_by_D": {str(D): sum(1 for p in plans if p[0] == D) for D in Ds},
           "hbm_budget_bytes": budget, "peak_pool_bytes": peak, "jumps": jumps,
           "wall_s": round(wall, 3), "trajectories_per_s": round(n / wall, 3),
           "trajectory_steps_per_s": round(n * steps / wall, 2),
           "sequential_fp64_trajectories_per_s": round(n / t64, 4),
           "sequential_fp32_trajectories_per_s": round(n / t32, 4),
           "speedup_vs_sequential_fp64": round(t64 / wall, 2),
           "speedup_vs_sequential_fp32": round(t32 / wall, 2),
           "achieved_gemm_tflops": round(total_flops / wall / 1e12, 2),
           "mfma_utilisation_fp32": round(total_flops / wall / 1e12 / PEAK["fp32"], 4),
           "note": "wall time includes building each pool and capturing its step graph; utilisation is "
                   "the reference's contraction flops over the leg's wall time against the FP32 MFMA peak"}
    print(f"ragged pools: {out}", file=sys.stderr, flush=True)
    return out


def batch_and_serve(seed):
    """Batched quantum-jump trajectories (concurrent per-stream solves vs the same
    solves one after another) and a Poisson request stream through the server."""
    from mpsamd import serving as S
    rng = np.random.default_rng(seed)
    n_traj, steps = 16, 10
    spec = [(int(rng.integers(40, 81)), int(rng.choice([16, 32])), int(rng.integers(1 << 30))) for _ in range(n_traj)]
    seq = [S.Request(i, "spin", N, D, steps, sd) for i, (N, D, sd) in enumerate(spec)]
    t_seq = S.run_batch(seq, sequential=True)
    print(f"sequential {n_traj} trajectories: {t_seq:.2f} s", file=sys.stderr, flush=True)
    con = [S.Request(i, "spin", N, D, steps, sd) for i, (N, D, sd) in enumerate(spec)]
    t_con = S.run_batch(con)
    print(f"concurrent {n_traj} trajectories: {t_con:.2f} s", file=sys.stderr, flush=True)
    same = max(float(np.abs(a.result - b.result).max()) for a, b in zip(seq, con))
    batched = {"trajectories": n_traj, "steps_each": steps, "N_range": [40, 80], "D_choices": [16, 32],
               "trajectories_per_s": round(n_traj / t_con, 3),
               "sequential_trajectories_per_s": round(n_traj / t_seq, 3),
               "speedup_vs_sequential": round(t_seq / t_con, 3), "max_abs_diff_vs_sequential": same}
    reqs = S.poisson_requests(40, rate=50.0, seed=seed, N_range=(20, 60), D_choices=(16, 32),
                              steps_range=(5, 20), rydberg_N_range=(10, 30), rydberg_D_choices=(16,))
    st = S.Server(hbm_budget_bytes=64 << 30, max_active=16).run(reqs)
    print(f"serving: {st}", file=sys.stderr, flush=True)
    serving = {"requests": st["requests"], "arrival_rate_req_s": 50.0,
               "mix": "70% spin trajectories (d=2, N 20-60, D 16/32), 30% Rydberg density matrices "
                      "(d=9, N 10-30, D 16), 5-20 steps each",
               "throughput_req_s": round(st["throughput_req_s"], 3),
               "p50_latency_s": round(st["p50_latency_s"], 4), "p99_latency_s": round(st["p99_latency_s"], 4)}
    return batched, serving


def pooled_serving_leg(seed, n_burst=480, n=320, rate=30.0, systems=4, N=60, D=32, slots=256, slots_poisson=128):
    """Trajectory requests over a few chains of one shape, continuously batched
    in a lockstep pool (one batched context, each slot with its own chain's
    MPOs, requests join and leave slots between ticks; the pool is built
    before the clock starts).  A burst (every
    request at t = 0) measures throughput; a Poisson stream below that
    capacity measures latency, on a pool half as wide (a shorter tick)."""
    from mpsamd import serving as S
    sy = [(N, D, seed + i) for i in range(systems)]
    out = {"chains": systems, "slots": slots,
           "mix": f"quantum-jump trajectories of {systems} driven chains (d=2, N={N}, D={D}), 10-40 steps each"}
    burst = S.poisson_trajectory_requests(n_burst, rate=1e12, seed=seed, systems=sy, steps_range=(10, 40))
    st = S.PooledServer(slots=slots, dt=0.01).run(burst, warm=[(N, D)])
    print(f"pooled serving, burst: {st}", file=sys.stderr, flush=True)
    out["burst"] = {"requests": st["requests"], "throughput_req_s": round(st["throughput_req_s"], 3),
                    "p50_latency_s": round(st["p50_latency_s"], 4), "p99_latency_s": round(st["p99_latency_s"], 4)}
    reqs = S.poisson_trajectory_requests(n, rate=rate, seed=seed + 1, systems=sy, steps_range=(10, 40))
    st = S.PooledServer(slots=slots_poisson, dt=0.01).run(reqs, warm=[(N, D)])
    print(f"pooled serving, Poisson {rate}/s: {st}", file=sys.stderr, flush=True)
    out["poisson"] = {"requests": st["requests"], "arrival_rate_req_s": rate, "slots": slots_poisson,
                      "throughput_req_s": round(st["throughput_req_s"], 3),
                      "p50_latency_s": round(st["p50_latency_s"], 4), "p99_latency_s": round(st["p99_latency_s"], 4)}
    # one stream of both kinds: spin trajectories and d=9 density matrices,
    # each kind continuously batched in its own lockstep pool
    mix_rate, dm_shape = 50.0, (40, 32)
    mix_sy = sy + [("rydberg", *dm_shape, seed + 10 + i) for i in range(2)]
    reqs = S.poisson_trajectory_requests(300, rate=mix_rate, seed=seed + 2, systems=mix_sy, steps_range=(10, 40))
    st = S.PooledServer(slots=slots_poisson, dm_slots=32, hbm_budget_bytes=64 << 30, dt=0.01).run(
        reqs, warm=[(N, D), ("rydberg", *dm_shape)])
    print(f"pooled serving, mixed Poisson {mix_rate}/s: {st}", file=sys.stderr, flush=True)
    kinds = {}
    for k in ("spin", "rydberg"):
        lat = np.array([r.done - r.arrival for r in reqs if r.kind == k])
        kinds[k] = {"requests": int(lat.size), "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                    "p99_latency_s": round(float(np.percentile(lat, 99)), 4)}
    out["mixed_poisson"] = {
        "mix": f"{len(sy)} spin chains (d=2, N={N}, D={D}, {slots_poisson} slots) and 2 Rydberg density-matrix "
               f"chains (d=9, N={dm_shape[0]}, D={dm_shape[1]}, 32 slots), drawn uniformly; 10-40 steps each",
        "requests": st["requests"], "arrival_rate_req_s": mix_rate,
        "throughput_req_s": round(st["throughput_req_s"], 3),
        "p50_latency_s": round(st["p50_latency_s"], 4), "p99_latency_s": round(st["p99_latency_s"], 4),
        "hbm_budget_bytes": st["hbm_budget_bytes"], "peak_pool_bytes": st["peak_pool_bytes"],
        "by_kind": kinds}
    out["three_forms"] = three_forms_leg(seed)
    return out


def three_forms_leg(seed, n=240, rate=50.0, budget=96 << 30, sample_per_kind=4):
    """A Poisson stream drawing from the three evolution forms: quantum-jump
    spin trajectories (d = 2, FP64), Rydberg density matrices (d = 9, FP64) and
    the same density matrices in the precision form (FP32 contractions + two
    FP64 refinement half-sweeps), over several (N, D) shapes, each (form,
    shape) continuously batched in its own pool, pools admitted against an
    HBM budget.  A seeded sample of each form's responses is solved again
    alone on a plain context in FP64 (the path the parity tests pin to the
    oracle) and compared."""
    from mpsamd import serving as S
    systems = ([("spin", Nn, Dd, seed + 20 + 2 * Nn + Dd) for Nn in (20, 60, 100) for Dd in (32, 64)]
               + [("rydberg", Nn, 32, seed + 40 + Nn) for Nn in (40, 60)]
               + [("rydberg_p", 40, 32, seed + 80)])
    reqs = S.poisson_trajectory_requests(n, rate=rate, seed=seed + 3, systems=systems, steps_range=(10, 30))
    st = S.PooledServer(slots=64, dm_slots=32, max_pools=len(systems), hbm_budget_bytes=budget, dt=0.01).run(reqs)
    print(f"three-form serving: {st}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(seed + 4)
    kinds, acc = {}, {}
    for k in ("spin", "rydberg", "rydberg_p"):
        mine = [r for r in reqs if r.kind == k]
        lat = np.array([r.done - r.arrival for r in mine])
        kinds[k] = {"requests": len(mine), "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                    "p99_latency_s": round(float(np.percentile(lat, 99)), 4)}
        errs = []
        for i in rng.choice(len(mine), size=min(sample_per_kind, len(mine)), replace=False):
            r = mine[int(i)]
            ref = S.TrajRequest(r.rid, r.N, r.D, r.pos_seed, r.seed, r.steps,
                                kind="rydberg" if k == "rydberg_p" else k)
            S.solve_alone(ref, dt=0.01) if k == "spin" else S.solve_alone(ref)
            errs.append(float(np.abs(r.result - ref.result).max()))
        acc[k] = {"sampled": len(errs), "max_abs_err_vs_fp64_alone": max(errs),
                  "median_abs_err_vs_fp64_alone": float(np.median(errs))}
    out = {"mix": "spin trajectories N in {20, 60, 100} x D in {32, 64} (64 slots each); Rydberg density "
                  "matrices N in {40, 60}, D = 32 (32 slots), FP64 and (N = 40) precision form; 10-30 steps "
                  "each, chains drawn uniformly",
           "requests": st["requests"], "arrival_rate_req_s": rate, "throughput_req_s": round(st["throughput_req_s"], 3),
           "p50_latency_s": round(st["p50_latency_s"], 4), "p99_latency_s": round(st["p99_latency_s"], 4),
           "pools_opened": st["pools_opened"], "hbm_budget_bytes": st["hbm_budget_bytes"],
           "peak_pool_bytes": st["peak_pool_bytes"], "by_kind": kinds, "accuracy": acc,
           "accuracy_note": "results: spin [<sigma_z^j>..., last p_coh], density matrix [trace, Rydberg population];"
                            " each sampled response against the same request alone in FP64"}
    print(f"three-form serving: {out}", file=sys.stderr, flush=True)
    return out


def qr_panels_leg(seed, m=1152, n=128, reps=60, ranks=(128, 60, 20), tol=1e-13):
    """leftorth! (orthonormalisation.jl:164-171) on one site whose panel is the
    headline workload's bulk shape (D*d x D = 1152 x 128), full rank and of
    rank r plus rounding-level noise under the 1e-13 cutoff: the QR time that
    the RK4 step pays about 400 times.  Each factorization is timed from the
    host around the call and a synchronize (launch and sync included)."""
    import mpsamd as M
    ctx = M.Context(0)
    rng = np.random.default_rng(seed)
    st = M.MPS(1, [1], [m, n], ctx)
    out = {"shape": [m, n], "reps": reps, "chain_path": os.environ.get("MPS_QR_CHAIN", "1") != "0"}
    for r in ranks:
        a = (rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))) @ \
            (rng.standard_normal((r, n)) + 1j * rng.standard_normal((r, n)))
        if r < n:
            a = a + 1e-17 * (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))
        ctx.set_qr_rank_tol(tol if r < n else 0.0)
        ts = []
        for i in range(reps):
            st.upload([a.reshape(m, 1, n, order="F")])
            ctx.synchronize()
            t0 = time.perf_counter()
            M.leftorth_(st, 0)
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        out[f"rank{r}_us"] = round(float(np.median(ts[5:])) * 1e6, 1)
    ctx.set_qr_rank_tol(0.0)
    return out


def qr_batched_leg(seed, cases=((2, 256, 16), (9, 128, 16), (2, 256, 64), (9, 32, 64)), reps=20, tol=1e-13):
    """leftorth! (orthonormalisation.jl:164-171) of one site on every copy of
    a lockstep batch at once (a pool's QR): nb copies of a (D, d, D) site,
    panel D*d x D, full rank and of rank D/5 plus rounding-level noise under
    the 1e-13 cutoff.  Each call is timed from the host around the call and a
    synchronize; the context's QR form follows MPS_QR_CHAIN (single launch
    unless it is 0)."""
    import mpsamd as M
    from mpsamd.trajectory import _state_bytes
    rng = np.random.default_rng(seed)
    out = {"reps": reps, "chain_path": os.environ.get("MPS_QR_CHAIN", "1") != "0", "cases": []}
    for d, D, nb in cases:
        dims = [D, D, D]
        ctx = M.Context.batched(nb, _state_bytes(dims, d) + 4096, M.batch_workspace_bytes(nb, D, d, 1), 0)
        comp = ctx.companion()
        ctx.set_qr_rank_tol(tol)
        g = M.MPS(2, d, dims, ctx)
        views = [g.view(b, comp) for b in range(nb)]
        row = {"d": d, "D": D, "copies": nb, "panel": [D * d, D]}
        for label, r in (("full", D), ("rank_D/5", D // 5)):
            sites = []
            for _ in range(nb):
                a = (rng.standard_normal((D * d, r)) + 1j * rng.standard_normal((D * d, r))) @ \
                    (rng.standard_normal((r, D)) + 1j * rng.standard_normal((r, D)))
                if r < D:
                    a = a + 1e-17 * (rng.standard_normal((D * d, D)) + 1j * rng.standard_normal((D * d, D)))
                s1 = rng.standard_normal((D, d, D)) + 1j * rng.standard_normal((D, d, D))
                sites.append([a.reshape((D, d, D), order="F"), s1])
            ts = []
            for i in range(reps):
                for v, st in zip(views, sites):
                    v.upload(st)
                ctx.synchronize()
                t0 = time.perf_counter()
                M.leftorth_(g, 0)
                ctx.synchronize()
                ts.append(time.perf_counter() - t0)
            row[f"{label}_us"] = round(float(np.median(ts[3:])) * 1e6, 1)
        out["cases"].append(row)
        print(f"qr batched: {row}", file=sys.stderr, flush=True)
        for v in views:
            v.__del__()
        g.__del__()
        comp.close()
        ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=100)
    ap.add_argument("--D", type=int, default=128)
    ap.add_argument("--n-exp", type=int, default=4)
    ap.add_argument("--dt", type=float, default=0.005)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 and FP32+refine forms")
    ap.add_argument("--no-serve", action="store_true", help="skip the batched-trajectory and serving legs")
    ap.add_argument("--gemm-counters", choices=["fp64", "fp32"], default=None,
                    help="only the calibration copy and the dominant GEMM launches (for rocprofv3 --pmc runs)")
    ap.add_argument("--three-forms-only", action="store_true",
                    help="run only the serving leg over the three evolution forms")
    ap.add_argument("--ragged-only", action="store_true",
                    help="run only the uneven-shape ragged-pool leg (512 trajectories, D up to 256)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch kernels eagerly instead of replaying the captured RK4 step")
    ap.add_argument("--timing-pass-only", action="store_true",
                    help="run only the event-timed single-stream pass (for a kernel trace whose per-launch "
                         "durations compare with the roofline figures); value is that pass's rate")
    ap.add_argument("--timing-prec", choices=["fp64", "fp32"], default="fp64",
                    help="precision of the event-timed profiling pass (fp32: the precision form's operator "
                         "contractions; for a kernel trace of that form with --timing-pass-only)")
    ap.add_argument("--qr-panels", action="store_true",
                    help="only time leftorth! on 1152 x 128 panels of full and low rank (MPS_QR_CHAIN=0 in the "
                         "environment: the launch-per-panel factorization)")
    ap.add_argument("--qr-batched", action="store_true",
                    help="only time leftorth! on every copy of lockstep batches (MPS_QR_CHAIN=0 in the "
                         "environment: the launch-per-panel form)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run a device copy of this many MiB first (HBM counter calibration)")
    args = ap.parse_args()
    if args.timing_pass_only:
        args.no_fp32 = args.no_serve = args.no_cpu = True
        args.no_profile = False
        args.warmup = 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N>1 is not supported "
                                   "(trajectories/requests would shard across GPUs with no collectives)"}))
        sys.exit(2)

    if args.gemm_counters:
        print(json.dumps({"gemm_counters": gemm_counters(args.gemm_counters)}))
        return
    if args.qr_panels:
        print(json.dumps({"qr_panels": qr_panels_leg(args.seed)}))
        return
    if args.qr_batched:
        print(json.dumps({"qr_batched": qr_batched_leg(args.seed)}))
        return
    import mpsamd as M
    if args.three_forms_only:
        print(json.dumps({"three_forms": three_forms_leg(args.seed)}))
        return
    if args.ragged_only:
        print(json.dumps({"ragged_pools": ragged_pools_leg(args.seed)}))
        return
    ctx = M.Context(0)
    if args.no_graphs:
        ctx.set_graphs(False)
    if args.calib_copy_mib > 0:
        # one-site states of side n: the library's copy kernel moves n*n*16 B each way
        n = int((args.calib_copy_mib * (1 << 20) // 16) ** 0.5)
        src, dst = M.MPS(1, [1], [n, n], ctx), M.MPS(1, [1], [n, n], ctx)
        for _ in range(2):
            dst.copy_from(src)
        ctx.synchronize()
        print(f"calibration copy: {n * n * 16} bytes read and written per launch", file=sys.stderr, flush=True)
        del src, dst
    p, z, L, rho, rhos, envs, envop, dims = setup(args.N, args.D, args.n_exp, args.dt, args.seed, ctx)
    dt = args.dt
    rho0 = rho.to_host()

    def step():
        M.RK4stp_apply_H_half_(dt, L, rho, rhos[0], rhos[1], rhos[2], *envs, envop)

    for i in range(args.warmup):
        step()
        print(f"warmup step {i} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(0 if args.timing_pass_only else args.steps):
        step()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    sweeps = args.steps / elapsed if not args.timing_pass_only else float("nan")
    print(f"timed {args.steps} steps: {elapsed:.3f} s", file=sys.stderr, flush=True)

    from mpsamd import models as MD
    idm = M.MPS.from_host(MD.identity_vec_mps(args.N), ctx)
    nem = M.MPS.from_host(MD.sum_local_vec_mps(args.N, MD.HEE3), ctx)
    nrm = M.MPS.from_host(MD.sum_local_vec_mps(args.N, MD.HSS3), ctx)
    pem = M.MPS.from_host(MD.pair_sum_vec_mps(args.N, MD.HEE3), ctx)
    # the reference's photon observables (rydberg_density_matrix_gpu.jl:135-151,
    # 237-245): I_r = tr(E_R^dag E_R rho), I2_r, and g2 = I2_r / I_r^2
    from mpsamd import observables as OB
    ph = {k: M.MPS.from_host(v, ctx) for k, v in OB.photon_observables(z).items()}

    def observables(r):
        # total populations, the pair correlation of the excited level
        # 2 sum_{i<j} <n_i n_j> / (sum_i <n_i>)^2, and the light's g2 (None on a
        # state with no excitation, e.g. the ground state before any step)
        tr_ = M.scal_prod_no_conj(idm, r)
        n_e = M.scal_prod_no_conj(nem, r) / tr_
        pairs = M.scal_prod_no_conj(pem, r) / tr_
        i_r = M.scal_prod_no_conj(ph["IR"], r) / tr_
        i2_r = M.scal_prod_no_conj(ph["I2R"], r) / tr_
        return {"trace": tr_, "n_e": n_e, "n_r": M.scal_prod_no_conj(nrm, r) / tr_,
                "g2_e": (2.0 * pairs / n_e ** 2).real if n_e != 0 else None,
                "I_r": i_r.real, "I2_r": i2_r.real, "g2_r": OB.g2(i_r.real, i2_r.real)}

    tr = M.scal_prod_no_conj(idm, rho)
    ref_obs = observables(rho)

    # Precision forms: the same start state and number of steps as the FP64 run.
    #   fp32:        operator contractions in complex FP32 (3M) on MFMA; QR, norms,
    #                overlap terms FP64
    #   fp32_refine: the same, and every RK4 call refits its last stage with two
    #                FP64 half-sweeps (mps_ctx_set_refine_sweeps)
    # Errors are against the FP64 run's state: totals of the excited and Rydberg
    # populations, the largest per-site error of <sigma_z> = <n_e> - <n_g>, and
    # 1 - fidelity.
    prec_forms = {}
    if not args.no_fp32:
        def fidelity(a, b):
            ab = abs(M.scal_prod(a, b))
            return ab * ab / (abs(M.scal_prod(a, a)) * abs(M.scal_prod(b, b)))

        sz = [M.MPS.from_host(MD.local_population_vec_mps(args.N, j, MD.HEE3 - MD.HGG3), ctx)
              for j in range(args.N)]

        def sigma_z(r):
            t = M.scal_prod_no_conj(idm, r)
            return np.array([(M.scal_prod_no_conj(m, r) / t).real for m in sz])

        ref_sz = sigma_z(rho)
        rho64 = rho.similar().copy_from(rho)
        for form, nref in (("fp32", 0), ("fp32_refine", 2)):
            rho.upload(rho0)
            for r in rhos:
                r.copy_from(rho)
            ctx.set_precision("fp32")
            ctx.set_refine_sweeps(nref)
            for i in range(args.warmup):
                step()
            ctx.synchronize()
            t0 = time.perf_counter()
            for i in range(args.steps):
                step()
            ctx.synchronize()
            el = time.perf_counter() - t0
            ctx.set_precision("fp64")
            ctx.set_refine_sweeps(0)
            ob = observables(rho)
            prec_forms[form] = {
                "value": round(args.steps / el, 4), "ms_per_step": round(1e3 * el / args.steps, 3),
                "speedup_vs_fp64": round((args.steps / el) / sweeps, 3),
                "refine_half_sweeps": nref,
                "err_n_e": abs(ob["n_e"] - ref_obs["n_e"]), "err_n_r": abs(ob["n_r"] - ref_obs["n_r"]),
                "err_sigma_z_max_site": float(np.abs(sigma_z(rho) - ref_sz).max()),
                "err_trace": abs(ob["trace"] - ref_obs["trace"]),
                "rel_err_pair_corr_e": abs(ob["g2_e"] - ref_obs["g2_e"]) / abs(ref_obs["g2_e"]),
                "rel_err_I_r": abs(ob["I_r"] - ref_obs["I_r"]) / abs(ref_obs["I_r"]),
                "rel_err_g2_r": abs(ob["g2_r"] - ref_obs["g2_r"]) / abs(ref_obs["g2_r"]),
                "infidelity": 1.0 - fidelity(rho, rho64)}
            print(f"{form}: {prec_forms[form]}", file=sys.stderr, flush=True)
        rho.copy_from(rho64)
        del rho64

    roof = None
    prof = None
    if not args.no_profile:
        ctx.set_precision(args.timing_prec)
        ctx.set_profiling(True)
        tp = time.perf_counter()
        for i in range(args.steps):
            step()
        st = ctx.kernel_stats()
        ctx.set_profiling(False)
        ctx.set_precision("fp64")
        if args.timing_pass_only:
            sweeps = args.steps / (time.perf_counter() - tp)
        achieved = st["gemm_flops"] / (st["gemm_ms"] * 1e-3) / 1e12
        dl = DOMINANT_LAUNCH
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK[args.timing_prec],
                "unit": "TFLOP/s", "frac": round(achieved / PEAK[args.timing_prec], 4),
                "traffic": dl["hbm_read_bytes"] + dl["hbm_write_bytes"],
                "traffic_scope": f"HBM bytes of one {dl['kernel']} launch of shape {dl['shape']} "
                                 f"(rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated); algorithmic "
                                 f"{dl['algorithmic_bytes']:.4g} B; MFMA busy {dl['mfma_busy']}",
                "kernel": ("zgemm_f64_kernel" if args.timing_prec == "fp64" else "zgemm_f32_3m_kernel")
                          + " (+ split-K reduce)",
                "launches": st["gemm_launches"], "avg_launch_us": round(st["gemm_ms"] * 1e3 / max(st["gemm_launches"], 1), 2),
                # the kernels execute the 3M product (6 real flops per complex MAC against the
                # 8 counted in `achieved`): the MFMA units' share of their peak
                "frac_3m": round(achieved * 0.75 / PEAK[args.timing_prec], 4)}
        if args.timing_prec != "fp64":   # the counter figures are the FP64 launch's
            roof["traffic"] = None
            roof["traffic_scope"] = "not measured for this precision"
        prof = {"gemm_ms_per_step": st["gemm_ms"] / args.steps, "qr_ms_per_step": st["qr_ms"] / args.steps,
                "qr_launches_per_step": st["qr_launches"] / args.steps,
                "launches_checked_in_process": st["launches_checked"],
                "gemm_gflop_per_step": st["gemm_flops"] / args.steps / 1e9}
        print(f"profile: {prof}", file=sys.stderr, flush=True)

    batched = serving = None
    if not args.no_serve:
        uneven, per_solve = batch_and_serve(args.seed)
        batched = {"lockstep_ensemble": ensemble_leg(args.seed), "concurrent_uneven": uneven,
                   "density_matrix_pool": density_matrix_pool_leg(args.seed),
                   "uneven_pools": uneven_pools_leg(args.seed),
                   "ragged_pools": ragged_pools_leg(args.seed)}
        serving = {"lockstep_pools": pooled_serving_leg(args.seed), "per_solve_streams": per_solve}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.N, args.D, args.n_exp, args.dt, args.seed)

    out = {
        "metric": METRIC,
        "value": round(sweeps, 4),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / sweeps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Rydberg ladder density matrix, vectorised Lindbladian, N={args.N}, d=9, "
                               f"w={p.w} ({args.n_exp} exponentials, soft core {p.v_max:g}), D_max={args.D}, "
                               f"dt={dt}, RK4stp_apply_H_half! from mpsgroundstate, QR rank cutoff {QR_RANK_TOL:g}",
                   "seed": args.seed, "model_gflop_per_step": round(model_flops(dims, 9, p.w) / 1e9, 1),
                   "trace_rho": [round(tr.real, 12), round(tr.imag, 12)],
                   "observables": {"n_e": ref_obs["n_e"].real, "n_r": ref_obs["n_r"].real,
                                   "pair_corr_e": ref_obs["g2_e"], "I_r": ref_obs["I_r"],
                                   "I2_r": ref_obs["I2_r"], "g2_r": ref_obs["g2_r"]}},
        "roofline": roof,
        "cpu_baseline": cpu,
        "profile": prof,
        "precision_forms": prec_forms,
        "batched_trajectories": batched,
        "serving": serving,
    }
    if args.timing_pass_only:
        out["mode"] = "timing pass only (single stream, every GEMM bracketed by events)"
        out["timing_precision"] = args.timing_prec
    print(json.dumps(out))


if __name__ == "__main__":
    main()
