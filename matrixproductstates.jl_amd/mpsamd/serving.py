"""Batched trajectories and a continuously batched evolution server.

Every solve (a quantum-jump trajectory or a density-matrix evolution) owns a
library context, i.e. its own HIP stream, workspace arena and captured RK4
graph, and keeps all of its tensors resident in HBM until it finishes.  One
scheduler tick first queues one step of every active solve (graph launches,
no host wait), then collects the host-side scalars (norms, jump draws) solve
by solve.  Independent solves therefore overlap on the device: the QR panels
and small GEMMs of one chain occupy a few CUs, and other chains fill the rest.

HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default); a
process that serves many solves at once should raise it (bench.py sets 16)
before its first HIP call.

The server admits requests as they arrive (Poisson, seeded), while the HBM
footprint of the active set stays under a budget and a slot is free, and
reports throughput and p50/p99 latency.  A request's result does not depend
on what else is running: each solve has its own stream and workspace, and
every kernel reduces in a fixed order.

Requests of one kind and shape (spin trajectories or density matrices; chain
length N, bond dimension D; the chains themselves may differ) are batched in
lockstep instead (``PooledServer``): a pool holds ``slots`` copies on one batched context, each
with its own MPOs, every tick advances all of them with one launch per
kernel, and a request joins a free slot as soon as it arrives and leaves it as
soon as it is done (continuous batching).  Each copy's arithmetic reads only
its own data, in a fixed order, so a response is the same whichever slot it
took and whatever else shared the pool.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import core as M
from . import models as MD
from .trajectory import SpinEnsemble, SpinTrajectory, slab_bytes, slab_sizes, spin_ensemble_bytes


class RydbergEvolution:
    """Density-matrix RK4 evolution of the Rydberg ladder (rydberg_density_matrix_gpu.jl:211-222)
    with the same launch/finish interface as ``SpinTrajectory``."""

    def __init__(self, N, D, seed, dt=0.005, n_exp=2, ctx: M.Context | None = None):
        self.N, self.D, self.dt = N, D, dt
        self.ctx = ctx if ctx is not None else M.default_context()
        p = MD.RydbergParams(n_exp=n_exp)
        self.z = MD.positions(N, float(N), seed)
        self.L = [M.MPO(h, self.ctx) for h in MD.rydberg_rk4_mpos(p, self.z, dt)]
        self.rho, self.dims = M.mpsgroundstate(N, D, 9, self.ctx)
        self.r = [self.rho.similar() for _ in range(3)]
        for s in self.r:
            s.copy_from(self.rho)
        self.env = [M.build_env(self.dims, self.dims, ctx=self.ctx) for _ in range(3)]
        self.envop = M.build_env(self.dims, self.dims, [1] + [p.w] * (N - 1) + [1], ctx=self.ctx)
        self.idm = M.MPS.from_host(MD.identity_vec_mps(N), self.ctx)
        self.nrm = M.MPS.from_host(MD.sum_local_vec_mps(N, MD.HSS3), self.ctx)
        self.w = p.w

    def launch_coherent(self):
        M.RK4stp_apply_H_half_(self.dt, self.L, self.rho, *self.r, *self.env, self.envop)

    def finish_step(self):
        self.ctx.synchronize()
        return 1.0

    def result(self):
        tr = M.scal_prod_no_conj(self.idm, self.rho)
        return np.array([tr.real, (M.scal_prod_no_conj(self.nrm, self.rho) / tr).real])

    def checkpoint(self, path, **meta):
        """Snapshot of rho between steps (the reference's _temp.mat role,
        rydberg_density_matrix_gpu.jl:251-255), in the mpsamd.checkpoint format."""
        from .checkpoint import save_state
        save_state(path, self.rho, kind="rydberg_evolution", N=self.N, D=self.D, dt=self.dt, **meta)

    def resume(self, path):
        """Continue from a snapshot (rydberg_pollution_density_matrix_gpu_restart.jl:
        151-161): later steps equal the uninterrupted run's bit for bit."""
        from .checkpoint import load_state
        _, meta = load_state(path, into=self.rho)
        if meta.get("kind") != "rydberg_evolution" or meta["N"] != self.N or meta["D"] != self.D:
            raise ValueError(f"{path}: not a snapshot of this evolution's shape")
        return meta

    def footprint_bytes(self):
        site = sum(self.dims[j] * 9 * self.dims[j + 1] for j in range(self.N))
        ovl = sum(self.dims[j] ** 2 for j in range(self.N + 1))
        env_op = sum(self.dims[j] ** 2 * self.w for j in range(self.N + 1))
        return 16 * (4 * site + 3 * ovl + env_op + 3 * self.N * self.w * self.w * 81)


@dataclass
class Request:
    rid: int
    kind: str            # "spin" (quantum-jump trajectory, d = 2) or "rydberg" (density matrix, d = 9)
    N: int
    D: int
    steps: int
    seed: int
    arrival: float = 0.0
    start: float = 0.0
    done: float = 0.0
    result: np.ndarray | None = None
    footprint: int = 0
    solve: object = field(default=None, repr=False)
    left: int = 0


# Requests start from product states at truncated bond dimension, whose QR
# panels are rank-deficient; the rank cutoff makes their Householder
# completions exact instead of noise-seeded (a context per request, so the
# setting stays local).
QR_RANK_TOL = 1e-13


def make_solve(req: Request, ctx: M.Context):
    ctx.set_qr_rank_tol(QR_RANK_TOL)
    if req.kind == "spin":
        return SpinTrajectory(req.N, req.D, req.seed, dt=0.01, ctx=ctx)
    return RydbergEvolution(req.N, req.D, req.seed, ctx=ctx)


def estimate_footprint(req: Request) -> int:
    """HBM bytes of a request before it is built: states + environments + MPOs, complex FP64,
    at maximal bond dimension, plus the workspace arena (about 4 sites of w D^2 d)."""
    d = 2 if req.kind == "spin" else 9
    w = 4 if req.kind == "spin" else 10
    nst = 7 if req.kind == "spin" else 4
    D = req.D
    site = req.N * D * D * d
    return 16 * (nst * site + 4 * req.N * D * D * w + 4 * D * D * d * w)


def release(solve):
    """Destroy a solve's device objects, then its context (objects must go first:
    they reference the context)."""
    for v in list(vars(solve).values()):
        for o in (v if isinstance(v, (list, tuple)) else [v]):
            if isinstance(o, (M.MPS, M.MPO, M.Env)):
                o.__del__()
    solve.ctx.close()


def _result(solve):
    if isinstance(solve, SpinTrajectory):
        return np.concatenate([solve.sigma_z(), [solve.p_coh[-1] if solve.p_coh else 1.0]])
    return solve.result()


def snapshot_requests(reqs, directory):
    """Snapshot of in-flight requests between ticks: each request's solve state
    (``<directory>/req<rid>.npz``) and a manifest of the requests (steps left,
    kind, shape, seed) in ``<directory>/requests.json``."""
    import json
    import os
    os.makedirs(directory, exist_ok=True)
    man = []
    for r in reqs:
        path = os.path.join(directory, f"req{r.rid}.npz")
        r.solve.checkpoint(path)
        man.append({"rid": r.rid, "kind": r.kind, "N": r.N, "D": r.D, "steps": r.steps, "seed": r.seed,
                    "left": r.left, "arrival": r.arrival, "state": os.path.basename(path)})
    tmp = os.path.join(directory, "requests.json.tmp")
    with open(tmp, "w") as f:
        json.dump(man, f)
    os.replace(tmp, os.path.join(directory, "requests.json"))


def restore_requests(directory):
    """The requests of a snapshot, each with its solve rebuilt on a fresh
    context and resumed, ready for further ticks (``run_batch(..., resumed=True)``)."""
    import json
    import os
    with open(os.path.join(directory, "requests.json")) as f:
        man = json.load(f)
    out = []
    for m in man:
        r = Request(m["rid"], m["kind"], m["N"], m["D"], m["steps"], m["seed"], m["arrival"])
        r.solve = make_solve(r, M.Context(0))
        r.solve.resume(os.path.join(directory, m["state"]))
        r.left = m["left"]
        out.append(r)
    return out


def run_batch(reqs, sequential=False, max_ticks=None, resumed=False):
    """Run every request to completion (or for ``max_ticks`` ticks, leaving the
    unfinished ones in flight); returns wall seconds.  ``sequential`` solves
    them one after another on one context (the plain baseline); ``resumed``:
    the requests already carry their solves (``restore_requests``)."""
    t0 = time.perf_counter()
    if sequential:
        ctx = M.Context(0)
        for r in reqs:
            r.solve = make_solve(r, ctx)
            for _ in range(r.steps):
                r.solve.launch_coherent()
                r.solve.finish_step()
            r.result = _result(r.solve)
            for v in list(vars(r.solve).values()):
                for o in (v if isinstance(v, (list, tuple)) else [v]):
                    if isinstance(o, (M.MPS, M.MPO, M.Env)):
                        o.__del__()
            r.solve = None
        ctx.close()
        return time.perf_counter() - t0
    active = []
    for r in reqs:
        if not resumed:
            r.solve = make_solve(r, M.Context(0))
            r.left = r.steps
        active.append(r)
    ticks = 0
    while active and (max_ticks is None or ticks < max_ticks):
        for r in active:
            r.solve.launch_coherent()
        for r in active:
            r.solve.finish_step()
            r.left -= 1
        for r in [r for r in active if r.left == 0]:
            r.result = _result(r.solve)
            release(r.solve)
            r.solve = None
            active.remove(r)
        ticks += 1
    return time.perf_counter() - t0


class Server:
    """Continuous batching over a request stream.

    ``requests`` carry arrival times (seconds from start).  Each tick admits
    the arrived requests that fit (slots, HBM budget; first come, first
    served), advances every active request by one step, and retires the
    finished ones.  Latency = completion - arrival.
    """

    def __init__(self, hbm_budget_bytes: int, max_active: int = 16):
        self.budget = hbm_budget_bytes
        self.max_active = max_active

    def run(self, requests):
        pending = sorted(requests, key=lambda r: r.arrival)
        active, done = [], []
        used = 0
        t0 = time.perf_counter()
        busy = 0.0
        while pending or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now and len(active) < self.max_active:
                r = pending[0]
                r.footprint = estimate_footprint(r)
                if used + r.footprint > self.budget and active:
                    break
                pending.pop(0)
                r.start = now
                r.solve = make_solve(r, M.Context(0))
                r.left = r.steps
                used += r.footprint
                active.append(r)
            if not active:
                time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                continue
            tb = time.perf_counter()
            for r in active:
                r.solve.launch_coherent()
            for r in active:
                r.solve.finish_step()
                r.left -= 1
            busy += time.perf_counter() - tb
            for r in [r for r in active if r.left == 0]:
                r.result = _result(r.solve)
                r.done = time.perf_counter() - t0
                release(r.solve)
                r.solve = None
                used -= r.footprint
                active.remove(r)
                done.append(r)
        wall = time.perf_counter() - t0
        lat = np.array([r.done - r.arrival for r in done])
        return {"requests": len(done), "wall_s": wall, "busy_s": busy,
                "throughput_req_s": len(done) / wall,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99))}


def poisson_requests(n, rate, seed, spin_frac=0.7, N_range=(20, 60), D_choices=(16, 32), steps_range=(10, 40),
                     rydberg_N_range=(10, 30), rydberg_D_choices=(16, 32)):
    """Seeded request stream: exponential inter-arrival times at ``rate`` per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        if rng.random() < spin_frac:
            out.append(Request(i, "spin", int(rng.integers(*N_range, endpoint=True)), int(rng.choice(D_choices)),
                               int(rng.integers(*steps_range, endpoint=True)), int(rng.integers(1 << 30)), t))
        else:
            out.append(Request(i, "rydberg", int(rng.integers(*rydberg_N_range, endpoint=True)),
                               int(rng.choice(rydberg_D_choices)), int(rng.integers(*steps_range, endpoint=True)),
                               int(rng.integers(1 << 30)), t))
    return out


# ------------------------------------------------------ lockstep trajectory pools
@dataclass
class TrajRequest:
    """One quantum-jump trajectory of a system: (N, D, pos_seed) fixes the
    chain and its MPOs, ``seed`` the jump draws, ``steps`` its length."""
    rid: int
    N: int
    D: int
    pos_seed: int
    seed: int
    steps: int
    arrival: float = 0.0
    start: float = 0.0
    done: float = 0.0
    result: np.ndarray | None = None
    left: int = 0
    # "spin": quantum-jump trajectory (d = 2), result [<sigma_z^j>..., p_coh];
    # "rydberg": density-matrix evolution (d = 9, no draws: seed unused),
    # result [trace, Rydberg-level population / trace]; "rydberg_p": the same
    # evolution in the precision form (operator contractions in complex FP32,
    # two FP64 refinement half-sweeps of every step's last stage)
    kind: str = "spin"


class TrajectoryPool:
    """Continuous lockstep batching of same-shape trajectories.

    ``slots`` copies of an (N, D) chain live on one batched context
    (``SpinEnsemble``); every copy holds its own MPOs, so the slots may follow
    different chains of that shape.  A request takes a free slot (its copy
    set to the request's chain, reset to the initial product state, its draws
    to the request's seed); every tick advances all slots with one batched
    RK4 call and one batched norm, then runs the jump branch of the occupied
    slots; a finished request's slot is free for the next one at once.  A
    slot's arithmetic is independent of the other slots, so a response does
    not depend on what else runs.
    """

    def __init__(self, N, D, slots=64, pos_seed=0, **system):
        self.key = ("spin", N, D)
        self.ens = SpinEnsemble(N, D, pos_seed, [0] * slots, qr_rank_tol=QR_RANK_TOL, **system)
        self.req = [None] * slots

    def free_slots(self):
        return [b for b, r in enumerate(self.req) if r is None]

    def pick_slot(self, r: TrajRequest, free):
        """A free slot already on the request's chain if there is one (no MPO
        upload), else the first free slot."""
        for b in free:
            if self.ens.system[b] == r.pos_seed:
                return b
        return free[0]

    def admit(self, r: TrajRequest, b: int):
        self.ens.set_system(b, r.pos_seed)
        self.ens.reset(b, r.seed)
        self.req[b] = r
        r.left = r.steps

    def busy(self):
        return any(r is not None for r in self.req)

    def launch(self):
        """Queue the batched coherent step (no host wait)."""
        self.ens.launch()

    def finish(self):
        """Norms, jump branches and retirement; returns the finished requests."""
        active = [b for b, r in enumerate(self.req) if r is not None]
        self.ens.finish(active)
        out = []
        for b in active:
            r = self.req[b]
            r.left -= 1
            if r.left == 0:
                r.result = np.concatenate([self.ens.sigma_z(b), [self.ens.p_coh[b][-1]]])
                self.req[b] = None
                out.append(r)
        return out

    def close(self):
        self.ens.close()


class DensityMatrixPool:
    """Continuous lockstep batching of same-shape density-matrix evolutions.

    ``slots`` copies of the vectorised Rydberg-ladder density matrix (N sites,
    d = 9, bond D, from ``mpsgroundstate``) live on one batched context, each
    with its own three RK4 Lindblad MPOs, so the slots may follow chains of
    different positions.  Every tick advances all copies with one batched
    ``RK4stp_apply_H_half!`` (rydberg_density_matrix_gpu.jl:211-222, the
    10-argument form); a request takes a free slot (its MPOs set to the
    request's chain, its copy reset to the ground state) and, after its
    ``steps``, reads the trace and the Rydberg-level population of its copy through
    a view.  ``RydbergEvolution`` is the same evolution on a plain context.
    """

    def __init__(self, N, D, slots=32, pos_seed=0, dt=0.005, n_exp=2, device=0, precision="fp64", refine=0):
        self.key = ("rydberg" if (precision, refine) == ("fp64", 0) else "rydberg_p", N, D)
        self.N, self.D, self.dt = N, D, dt
        self.p = MD.RydbergParams(n_exp=n_exp)
        Ls = MD.rydberg_rk4_mpos(self.p, MD.positions(N, float(N), pos_seed), dt)
        dims = M._dims_rule(N, D, 9)
        # the refinement keeps two more states of rho's shape per copy
        copy_bytes, ws, w = slab_sizes(dims, 9, Ls, 4 + (2 if refine else 0), slots, device)
        self.hbm_bytes = slab_bytes(copy_bytes, ws, slots)
        self.ctx = M.Context.batched(slots, copy_bytes, ws, device)
        self.ctx.set_qr_rank_tol(QR_RANK_TOL)
        self.ctx.set_precision(precision)
        self.ctx.set_refine_sweeps(refine)
        self.comp = self.ctx.companion()
        self.comp.set_qr_rank_tol(QR_RANK_TOL)
        self.L = [M.MPO(h, self.ctx) for h in Ls]
        self.rho, self.dims = M.mpsgroundstate(N, D, 9, self.ctx)
        self.r = [self.rho.similar() for _ in range(3)]
        for s in self.r:
            s.copy_from(self.rho)
        self.env = [M.build_env(self.dims, self.dims, ctx=self.ctx) for _ in range(3)]
        self.envop = M.build_env(self.dims, self.dims, [1] + [w] * (N - 1) + [1], ctx=self.ctx)
        self._init, _ = M.mpsgroundstate(N, D, 9, self.comp)
        self.idm = M.MPS.from_host(MD.identity_vec_mps(N), self.comp)
        self.nrm = M.MPS.from_host(MD.sum_local_vec_mps(N, MD.HSS3), self.comp)
        self.rv = [self.rho.view(b, self.comp) for b in range(slots)]
        self.system = [pos_seed] * slots
        self.req = [None] * slots

    def set_system(self, b, pos_seed):
        """Copy b follows the chain with positions ``models.positions(N, N,
        pos_seed)`` from now on (its RK4 MPOs rewritten in place)."""
        if self.system[b] == pos_seed:
            return
        Ls = MD.rydberg_rk4_mpos(self.p, MD.positions(self.N, float(self.N), pos_seed), self.dt)
        for h_dev, h_host in zip(self.L, Ls):
            v = h_dev.view(b, self.comp)
            for j, site in enumerate(h_host):
                v.upload_site(j, site)
            v.__del__()
        self.system[b] = pos_seed

    def free_slots(self):
        return [b for b, r in enumerate(self.req) if r is None]

    def pick_slot(self, r: TrajRequest, free):
        for b in free:
            if self.system[b] == r.pos_seed:
                return b
        return free[0]

    def admit(self, r: TrajRequest, b: int):
        self.set_system(b, r.pos_seed)
        self.rv[b].copy_from(self._init)
        self.req[b] = r
        r.left = r.steps

    def busy(self):
        return any(r is not None for r in self.req)

    def launch(self):
        """Queue the batched RK4 step of every copy (no host wait)."""
        M.RK4stp_apply_H_half_(self.dt, self.L, self.rho, *self.r, *self.env, self.envop)

    def result(self, b):
        tr = M.scal_prod_no_conj(self.idm, self.rv[b])
        return np.array([tr.real, (M.scal_prod_no_conj(self.nrm, self.rv[b]) / tr).real])

    def finish(self):
        """Wait for the step; retire the requests that are done."""
        self.ctx.synchronize()
        out = []
        for b, r in enumerate(self.req):
            if r is None:
                continue
            r.left -= 1
            if r.left == 0:
                r.result = self.result(b)
                self.req[b] = None
                out.append(r)
        return out

    def close(self):
        for v in self.rv:
            v.__del__()
        for o in (self._init, self.idm, self.nrm, self.rho, *self.r, *self.env, self.envop, *self.L):
            o.__del__()
        self.comp.close()
        self.ctx.close()


def solve_alone(r: TrajRequest, **system):
    """The same request on its own plain context (the sequential baseline)."""
    ctx = M.Context(0)
    ctx.set_qr_rank_tol(QR_RANK_TOL)
    if r.kind == "rydberg_p":
        ctx.set_precision("fp32")
        ctx.set_refine_sweeps(2)
    if r.kind in ("rydberg", "rydberg_p"):
        tr = RydbergEvolution(r.N, r.D, r.pos_seed, ctx=ctx)
        for _ in range(r.steps):
            tr.launch_coherent()
            tr.finish_step()
        r.result = tr.result()
    else:
        tr = SpinTrajectory(r.N, r.D, r.seed, ctx=ctx, pos_seed=r.pos_seed, **system)
        for _ in range(r.steps):
            tr.step()
        r.result = np.concatenate([tr.sigma_z(), [tr.p_coh[-1]]])
    for v in list(vars(tr).values()):
        for o in (v if isinstance(v, (list, tuple)) else [v]):
            if isinstance(o, (M.MPS, M.MPO, M.Env)):
                o.__del__()
    ctx.close()
    return r


def density_pool_bytes(N, D, slots, n_exp=2, dt=0.005, device=0, refine=0):
    """HBM bytes of the slab of a ``DensityMatrixPool``, before building it
    (the MPO shapes do not depend on the positions)."""
    Ls = MD.rydberg_rk4_mpos(MD.RydbergParams(n_exp=n_exp), MD.positions(N, float(N), 0), dt)
    copy_bytes, ws, _ = slab_sizes(M._dims_rule(N, D, 9), 9, Ls, 4 + (2 if refine else 0), slots, device)
    return slab_bytes(copy_bytes, ws, slots)


class PooledServer:
    """Continuous batching of trajectory and density-matrix requests over
    per-shape pools.

    Requests arrive with arrival times; each tick admits the arrived requests
    into free slots of the pool of their kind and shape (kind, N, D)
    (``TrajectoryPool`` of ``slots`` copies for spin trajectories,
    ``DensityMatrixPool`` of ``dm_slots`` copies for density matrices; a pool is created,
    within ``max_pools``, when a shape first arrives), queues one batched step of
    every busy pool (each pool has its own stream), then collects the norms,
    runs the jump branches and retires finished requests; a pool's next step
    is queued as soon as its last one is collected.  A pool with nothing in
    flight is released when ``max_pools`` are open and another shape
    arrives.  With ``hbm_budget_bytes``, a pool opens only while the slabs of
    the open pools and its own (computed before it is built) stay within the
    budget: idle pools are released to make room, and otherwise its requests
    wait until room frees; a shape whose pool alone exceeds the budget is an
    error.  Latency = completion - arrival.  ``system`` (dt, Omega, Gamma,
    gam_eg, init ...) is shared by every spin pool.
    """

    def __init__(self, slots=64, max_pools=8, dm_slots=32, hbm_budget_bytes=None, **system):
        self.slots, self.max_pools, self.dm_slots, self.system = slots, max_pools, dm_slots, system
        self.budget = hbm_budget_bytes
        self._bytes = {}

    def pool_bytes(self, key):
        """HBM bytes of the slab of the pool for (kind, N, D)."""
        if key not in self._bytes:
            kind, N, D = key
            if kind in ("rydberg", "rydberg_p"):
                self._bytes[key] = density_pool_bytes(N, D, self.dm_slots, refine=2 if kind == "rydberg_p" else 0)
            else:
                self._bytes[key] = spin_ensemble_bytes(N, D, self.slots, self.system.get("dt", 0.01))
        return self._bytes[key]

    def _pool(self, key):
        kind, N, D = key
        if kind == "rydberg":
            return DensityMatrixPool(N, D, self.dm_slots)
        if kind == "rydberg_p":
            return DensityMatrixPool(N, D, self.dm_slots, precision="fp32", refine=2)
        return TrajectoryPool(N, D, self.slots, **self.system)

    def run(self, requests, warm=()):
        """Serve ``requests``; pools of the shapes in ``warm`` [(N, D) for
        spin trajectories, or (kind, N, D)] are built and stepped twice
        (capturing their step graphs) before the clock starts, as a running
        server keeps them."""
        pending = sorted(requests, key=lambda r: r.arrival)
        warm = [tuple(k) if len(k) == 3 else ("spin",) + tuple(k) for k in warm]
        if self.budget is not None:
            big = [k for k in {(r.kind, r.N, r.D) for r in requests} | set(warm) if self.pool_bytes(k) > self.budget]
            if big:
                raise ValueError(f"pools larger than the HBM budget: {big}")
        pools = {}
        acct = {"used": 0, "peak": 0, "opened": 0}

        def open_pool(key):
            pools[key] = pool = self._pool(key)
            if self.budget is not None:
                acct["used"] += self.pool_bytes(key)
                acct["peak"] = max(acct["peak"], acct["used"])
            acct["opened"] += 1
            return pool

        def close_pool(key):
            pools.pop(key).close()
            if self.budget is not None:
                acct["used"] -= self.pool_bytes(key)

        def fits(key):
            return len(pools) < self.max_pools and (
                self.budget is None or acct["used"] + self.pool_bytes(key) <= self.budget)

        for key in warm:
            if not fits(key):
                continue
            pool = open_pool(key)
            for _ in range(2):
                pool.launch()
                pool.finish()
        done = []
        inflight = set()
        t0 = time.perf_counter()

        def admit(now):
            keep = []
            for r in pending:
                if r.arrival > now:
                    keep.append(r)
                    continue
                key = (r.kind, r.N, r.D)
                pool = pools.get(key)
                if pool is None and not fits(key):
                    # make room: release pools with nothing in flight or queued
                    idle = [k for k, q in pools.items() if k not in inflight and not q.busy()]
                    while idle and not fits(key):
                        close_pool(idle.pop(0))
                if pool is None and fits(key):
                    pool = open_pool(key)
                # a pool with a step in flight takes requests only once that
                # step is collected (the step's norms belong to the old slots)
                free = pool.free_slots() if pool is not None and key not in inflight else []
                if not free:
                    keep.append(r)
                    continue
                r.start = now
                pool.admit(r, pool.pick_slot(r, free))
            return keep

        # Pools advance independently: a pool's next step is queued as soon as
        # its previous one is collected, so the other pools' queued steps keep
        # the GPU busy while the host handles one pool.  A pool opened by an
        # admission during a pass is queued on the next pass, so the loop runs
        # while any pool holds requests.
        while pending or inflight or any(q.busy() for q in pools.values()):
            pending = admit(time.perf_counter() - t0)
            for key, p in list(pools.items()):
                if key in inflight:
                    for r in p.finish():
                        r.done = time.perf_counter() - t0
                        done.append(r)
                    inflight.discard(key)
                    pending = admit(time.perf_counter() - t0)
                if p.busy():
                    p.launch()
                    inflight.add(key)
            if not inflight and pending:
                time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
        wall = time.perf_counter() - t0
        for p in pools.values():
            p.close()
        lat = np.array([r.done - r.arrival for r in done])
        out = {"requests": len(done), "wall_s": wall, "pools": len(pools), "pools_opened": acct["opened"],
               "throughput_req_s": len(done) / wall,
               "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99))}
        if self.budget is not None:
            out.update(hbm_budget_bytes=self.budget, peak_pool_bytes=acct["peak"])
        return out


def poisson_trajectory_requests(n, rate, seed, systems, steps_range=(10, 40)):
    """Seeded stream of requests over ``systems`` [(N, D, pos_seed) for spin
    trajectories, or (kind, N, D, pos_seed)]: exponential inter-arrival times
    at ``rate`` per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        s = systems[int(rng.integers(len(systems)))]
        kind, (N, D, ps) = (s[0], s[1:]) if len(s) == 4 else ("spin", s)
        out.append(TrajRequest(i, N, D, ps, int(rng.integers(1 << 30)),
                               int(rng.integers(steps_range[0], steps_range[1], endpoint=True)), t,
                               kind=kind))
    return out
