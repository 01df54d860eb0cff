"""Quantum-jump trajectories of the driven spin chain on the device.

Mirrors the time loop of spin_model_quantum_trajectory_rk4.jl:115-209:

* coherent step: ``RK4stp_apply_H_half!`` (11-argument form) into ``A_coh``,
  ``p_coh = normalize_lo!(A_coh)``;
* no jump when ``r1 < p_coh``: ``A <- A_coh``;
* otherwise the jump probabilities: right/left waveguide emission from
  ``compress_var_apply_H!`` with the bond-2 jump MPOs (one sweep) and
  ``normalize_lo!``, local emission from ``measure_excitations!`` times
  ``dt * gam_eg``; the cumulative table is normalised and ``r2`` picks the
  jump: ``apply_site_operator!(A[j], hge)`` + ``leftorthonormalizationQR!``,
  or ``A <- A_r`` / ``A <- A_l``.

The random draws come from ``numpy.random.default_rng(seed)`` (two uniforms
per step, r1 then r2), not Julia's dSFMT, so jump records are reproducible
here but not identical to the reference's.  Every tensor stays on the device;
the host sees only the scalars that decide the branch.
"""
from __future__ import annotations

import numpy as np

from . import core as M
from . import models as MD

HGE = MD.HGE2
HEE = MD.HEE2


class SpinTrajectory:
    """One trajectory of the driven two-level chain (bond-4 H_eff MPO).

    Positions are ``models.positions(N, L, seed)``; the initial state is the
    product state ``init`` on every site (default |g>) at maximal bond
    dimension ``D`` (``mpsproductstate``).
    """

    def __init__(self, N, D, seed, dt=0.01, Omega=0.5, Gamma=1.0, k=2 * np.pi, L=None, gam_eg=0.0,
                 init=(1.0, 0.0), ctx: M.Context | None = None, pos_seed=None):
        self.N, self.D, self.dt, self.seed, self.gam_eg = N, D, dt, seed, gam_eg
        self.ctx = ctx if ctx is not None else M.default_context()
        self.z = MD.positions(N, float(N if L is None else L), seed if pos_seed is None else pos_seed)
        self.H = [M.MPO(h, self.ctx) for h in MD.spin_rk4_mpos(self.z, 0.0, dt, Omega, Gamma, k)]
        jr, jl = MD.spin_jump_mpos(self.z, Gamma, k)
        self.jr, self.jl = M.MPO(jr, self.ctx), M.MPO(jl, self.ctx)
        self.A, self.dims = M.mpsproductstate(N, D, 2, np.asarray(init, dtype=complex), self.ctx)
        self.A_coh, self.A_r, self.A_l = (self.A.similar() for _ in range(3))
        self.A1, self.A2, self.A3 = (self.A.similar() for _ in range(3))
        for s in (self.A_coh, self.A_r, self.A_l, self.A1, self.A2, self.A3):
            s.copy_from(self.A)
        self.envop = M.build_env(self.dims, self.dims, [1] + [4] * (N - 1) + [1], ctx=self.ctx)
        self.envop_jump = M.build_env(self.dims, self.dims, [1] + [2] * (N - 1) + [1], ctx=self.ctx)
        self.env = [M.build_env(self.dims, self.dims, ctx=self.ctx) for _ in range(3)]
        self.rng = np.random.default_rng(seed)
        self.t = 0.0
        self.p_coh = []
        self.jumps = []          # (step, kind, site): kind 'r', 'l' or 'eg'

    # ------------------------------------------------------------ one step
    def launch_coherent(self):
        """Queue the coherent RK4 step on the context's stream (no host wait)."""
        M.RK4stp_apply_H_half_(self.dt, self.H, self.A_coh, self.A, self.A1, self.A2, self.A3,
                               *self.env, self.envop)

    def finish_step(self):
        """Read p_coh (synchronises the stream), draw r1/r2 and take the branch."""
        p = M.normalize_lo_(self.A_coh)
        self.p_coh.append(p)
        r1, r2 = self.rng.random(2)
        i = len(self.p_coh) - 1
        if r1 < p:
            self.A.copy_from(self.A_coh)
        else:
            M.compress_var_apply_H_(self.A_r, self.envop_jump, self.A, self.A, self.jr, 1)
            p_r = M.normalize_lo_(self.A_r)
            M.compress_var_apply_H_(self.A_l, self.envop_jump, self.A, self.A, self.jl, 1)
            p_l = M.normalize_lo_(self.A_l)
            epop = M.measure_excitations_(None, self.A, HEE)
            acc = jump_table(epop, p_r, p_l, self.dt, self.gam_eg)
            kind, j = pick_jump(acc, r2, self.N)
            if kind == "eg":
                M.apply_site_operator_(self.A, j, HGE)
                M.leftorthonormalizationQR_(self.A)
            elif kind == "r":
                self.A.copy_from(self.A_r)
            else:
                self.A.copy_from(self.A_l)
            self.jumps.append((i, kind, j))
        self.t += self.dt
        return p

    def step(self):
        self.launch_coherent()
        return self.finish_step()

    # ---------------------------------------------------- checkpoint / resume
    def checkpoint(self, path):
        """Snapshot between steps: the state A, the draws' generator state, the
        time, the p_coh record and the jump record (mpsamd.checkpoint format)."""
        from .checkpoint import save_state
        save_state(path, self.A, kind="spin_trajectory", N=self.N, D=self.D, seed=self.seed, t=self.t,
                   rng=self.rng.bit_generator.state, p_coh=self.p_coh, jumps=self.jumps)

    def resume(self, path):
        """Continue from a snapshot of a trajectory of the same system: the
        following steps equal the uninterrupted run's, bit for bit."""
        from .checkpoint import load_state
        _, meta = load_state(path, into=self.A)
        if meta.get("kind") != "spin_trajectory" or meta["N"] != self.N or meta["D"] != self.D:
            raise ValueError(f"{path}: not a snapshot of this trajectory's shape")
        self.rng.bit_generator.state = meta["rng"]
        self.t = float(meta["t"])
        self.p_coh = [float(x) for x in meta["p_coh"]]
        self.jumps = [(int(i), str(k), int(j)) for i, k, j in meta["jumps"]]

    def sigma_z(self):
        """<sigma_z^j> = 2 <n_e^j> - 1 (measure_excitations!, :220-231)."""
        return 2.0 * M.measure_excitations_(None, self.A, HEE) - 1.0

    def footprint_bytes(self):
        """HBM held by this trajectory: 7 states, 5 environments, 5 MPOs (complex FP64)."""
        site = sum(self.dims[j] * 2 * self.dims[j + 1] for j in range(self.N))
        ovl = sum(self.dims[j] ** 2 for j in range(self.N + 1))
        env_op = sum(self.dims[j] ** 2 * (1 if j in (0, self.N) else 4) for j in range(self.N + 1))
        env_jump = sum(self.dims[j] ** 2 * (1 if j in (0, self.N) else 2) for j in range(self.N + 1))
        mpo = self.N * (3 * 16 * 4 + 2 * 4 * 4)
        return 16 * (7 * site + 3 * ovl + env_op + env_jump + mpo)


def jump_table(epop, p_r, p_l, dt, gam_eg):
    """Normalised cumulative jump probabilities (spin_model_quantum_trajectory_rk4.jl:144-151):
    entries 1..N local emission, N+1 right, N+2 left."""
    N = len(epop)
    acc = np.zeros(N + 3)
    for j in range(N):
        acc[j + 1] = acc[j] + dt * gam_eg * epop[j]
    acc[N + 1] = acc[N] + dt * p_r
    acc[N + 2] = acc[N + 1] + dt * p_l
    return acc / acc[-1]


def pick_jump(acc, r2, N):
    """The branch taken by r2 (spin_model_quantum_trajectory_rk4.jl:154-177)."""
    if r2 < acc[N]:
        for j in range(N):
            if r2 < acc[j + 1]:
                return "eg", j
    if r2 < acc[N + 1]:
        return "r", -1
    return "l", -1


def _state_bytes(dims, d):
    return 16 * sum(dims[j] * d * dims[j + 1] for j in range(len(dims) - 1))


def _env_bytes(dims, w):
    N = len(dims) - 1
    tot = 0
    for j in range(N + 1):
        edge = j in (0, N)
        b = 1 if edge else dims[j]
        tot += b * b * (1 if edge else w)
    return 16 * tot


def _mpo_bytes(sites):
    return 16 * sum(s.size for s in sites)


def slab_sizes(dims, d, Hs, n_states, B, device=0):
    """(copy_bytes, ws_bytes, w) of a lockstep batch of B copies, each holding
    ``n_states`` states of ``dims`` with local dim d, three overlap
    environments, one operator environment and the MPOs ``Hs``."""
    w = max(h.shape[0] for hs in Hs for h in hs)
    copy_bytes = (n_states * (_state_bytes(dims, d) + 256) + 3 * (_env_bytes(dims, 1) + 256)
                  + _env_bytes(dims, w) + 256 + sum(_mpo_bytes(h) + 256 for h in Hs) + 256 * 64)
    ws = M.batch_workspace_bytes(B, max(dims), d, w, device)
    return copy_bytes, ws, w


def slab_bytes(copy_bytes, ws, B):
    """HBM bytes of the batched context's slab (the library's per-copy stride
    times B)."""
    ws = (ws + 255) & ~255
    return B * ((copy_bytes + ws + 4096 + 255) & ~255)


def spin_ensemble_bytes(N, D, B, dt=0.01, device=0):
    """HBM bytes of the slab of a ``SpinEnsemble`` of B copies of an (N, D)
    chain, before building it (the MPO shapes do not depend on the positions)."""
    Hs = MD.spin_rk4_mpos(MD.positions(N, float(N), 0), 0.0, dt)
    copy_bytes, ws, _ = slab_sizes(M._dims_rule(N, D, 2), 2, Hs, 5, B, device)
    return slab_bytes(copy_bytes, ws, B)


def ragged_ensemble_bytes(Ns, D, dt=0.01, device=0):
    """HBM bytes of the slab of a ``RaggedSpinEnsemble`` of chains of lengths
    Ns at bond dimension D, before building it."""
    nmax = max(Ns)
    Hs = MD.spin_rk4_mpos(MD.positions(nmax, float(nmax), 0), 0.0, dt)
    copy_bytes, ws, _ = slab_sizes(M._dims_rule(nmax, D, 2), 2, Hs, 5, len(Ns), device)
    return slab_bytes(copy_bytes, ws, len(Ns))


class SpinEnsemble:
    """A lockstep ensemble of B trajectories of one driven spin chain.

    All trajectories share the shapes (N, D, the MPO bonds); each copy holds
    its own MPOs, so copies may follow different chains (``set_system``; at
    creation every copy has positions from ``pos_seed``).  Trajectory b draws
    r1, r2 from ``numpy.random.default_rng(seeds[b])`` exactly as
    ``SpinTrajectory`` with ``seed = seeds[b]`` and the copy's ``pos_seed``
    does.  The coherent RK4 step and
    ``normalize_lo!`` of all B run as one batched call (every GEMM, QR panel
    and reduction launched once for the ensemble, on a batched context); the
    jump branch of a trajectory (spin_model_quantum_trajectory_rk4.jl:140-177)
    runs on views of its copy, on the companion context (same stream).
    """

    def __init__(self, N, D, pos_seed, seeds, dt=0.01, Omega=0.5, Gamma=1.0, k=2 * np.pi, L=None,
                 gam_eg=0.0, init=(1.0, 0.0), qr_rank_tol=0.0, device=0):
        self.N, self.D, self.dt, self.gam_eg = N, D, dt, gam_eg
        self._model = (float(N if L is None else L), Omega, Gamma, k)
        self.B = B = len(seeds)
        self.z = MD.positions(N, float(N if L is None else L), pos_seed)
        Hs = MD.spin_rk4_mpos(self.z, 0.0, dt, Omega, Gamma, k)
        jr, jl = MD.spin_jump_mpos(self.z, Gamma, k)
        dims = M._dims_rule(N, D, 2)
        copy_bytes, ws, w = slab_sizes(dims, 2, Hs, 5, B, device)
        self.hbm_bytes = slab_bytes(copy_bytes, ws, B)
        self.ctx = M.Context.batched(B, copy_bytes, ws, device)
        self.ctx.set_qr_rank_tol(qr_rank_tol)
        self.comp = self.ctx.companion()
        self.comp.set_qr_rank_tol(qr_rank_tol)
        self.H = [M.MPO(h, self.ctx) for h in Hs]
        self.jr, self.jl = M.MPO(jr, self.comp), M.MPO(jl, self.comp)
        self._jumps = {pos_seed: (self.jr, self.jl)}     # jump MPOs per chain (companion)
        self.system = [pos_seed] * B
        self.A, self.dims = M.mpsproductstate(N, D, 2, np.asarray(init, dtype=complex), self.ctx)
        self.A_coh, self.A1, self.A2, self.A3 = (self.A.similar() for _ in range(4))
        for s in (self.A_coh, self.A1, self.A2, self.A3):
            s.copy_from(self.A)
        self.envop = M.build_env(self.dims, self.dims, [1] + [w] * (N - 1) + [1], ctx=self.ctx)
        self.env = [M.build_env(self.dims, self.dims, ctx=self.ctx) for _ in range(3)]
        # jump scratch on the companion (one trajectory at a time)
        self.A_r = M.MPS(N, 2, self.dims, self.comp)
        self.A_l = M.MPS(N, 2, self.dims, self.comp)
        self.envop_jump = M.build_env(self.dims, self.dims, [1] + [2] * (N - 1) + [1], ctx=self.comp)
        self.Av = [self.A.view(b, self.comp) for b in range(B)]
        self.Acv = [self.A_coh.view(b, self.comp) for b in range(B)]
        self.rngs = [np.random.default_rng(s) for s in seeds]
        self._init_vec = init
        self.p_coh = [[] for _ in range(B)]
        self.jumps = [[] for _ in range(B)]
        self.t = 0.0

    def set_system(self, b, pos_seed):
        """Copy b follows the chain with positions ``models.positions(N, L,
        pos_seed)`` from now on (its RK4 MPOs rewritten in place)."""
        if self.system[b] == pos_seed:
            return
        L, Omega, Gamma, k = self._model
        z = MD.positions(self.N, L, pos_seed)
        for h_dev, h_host in zip(self.H, MD.spin_rk4_mpos(z, 0.0, self.dt, Omega, Gamma, k)):
            v = h_dev.view(b, self.comp)
            for j, site in enumerate(h_host):
                v.upload_site(j, site)
            v.__del__()
        if pos_seed not in self._jumps:
            jr, jl = MD.spin_jump_mpos(z, Gamma, k)
            self._jumps[pos_seed] = (M.MPO(jr, self.comp), M.MPO(jl, self.comp))
        self.system[b] = pos_seed

    def reset(self, b, seed):
        """Start a new trajectory in copy b: the initial product state, a fresh
        jump record and draws from default_rng(seed)."""
        if not hasattr(self, "_init"):
            self._init, _ = M.mpsproductstate(self.N, self.D, 2, np.asarray(self._init_vec, dtype=complex),
                                              self.comp)
        self.Av[b].copy_from(self._init)
        self.rngs[b] = np.random.default_rng(seed)
        self.p_coh[b] = []
        self.jumps[b] = []

    def launch(self):
        """Queue the batched coherent RK4 step of all B copies (no host wait)."""
        M.RK4stp_apply_H_half_(self.dt, self.H, self.A_coh, self.A, self.A1, self.A2, self.A3, *self.env,
                               self.envop)

    def step(self, active=None):
        """One step of every trajectory (the jump branch only for copies in
        ``active``, default all); returns the B coherent norms p_coh."""
        self.launch()
        return self.finish(active)

    def finish(self, active=None):
        """Batched ``normalize_lo!`` (waits for the stream), then the draws and
        jump branches of the copies in ``active``; copies outside it evolve
        coherently."""
        p = M.normalize_lo_batch(self.A_coh)
        for b in (range(self.B) if active is None else active):
            self.p_coh[b].append(float(p[b]))
            r1, r2 = self.rngs[b].random(2)
            i = len(self.p_coh[b]) - 1
            if r1 < p[b]:
                continue
            A, Ac = self.Av[b], self.Acv[b]
            jr, jl = self._jumps[self.system[b]]
            M.compress_var_apply_H_(self.A_r, self.envop_jump, A, A, jr, 1)
            p_r = M.normalize_lo_(self.A_r)
            M.compress_var_apply_H_(self.A_l, self.envop_jump, A, A, jl, 1)
            p_l = M.normalize_lo_(self.A_l)
            epop = M.measure_excitations_(None, A, HEE)
            acc = jump_table(epop, p_r, p_l, self.dt, self.gam_eg)
            kind, j = pick_jump(acc, r2, self.N)
            # the jump result replaces the candidate coherent state of copy b
            if kind == "eg":
                Ac.copy_from(A)
                M.apply_site_operator_(Ac, j, HGE)
                M.leftorthonormalizationQR_(Ac)
            elif kind == "r":
                Ac.copy_from(self.A_r)
            else:
                Ac.copy_from(self.A_l)
            self.jumps[b].append((i, kind, j))
        self.A.copy_from(self.A_coh)
        self.t += self.dt
        return p

    def sigma_z(self, b):
        """<sigma_z^j> of trajectory b."""
        return 2.0 * M.measure_excitations_(None, self.Av[b], HEE) - 1.0

    def close(self):
        """Release views, containers, then the two contexts."""
        for name in ("Av", "Acv"):
            for v in getattr(self, name):
                v.__del__()
        if hasattr(self, "_init"):
            self._init.__del__()
        for jr, jl in self._jumps.values():
            jr.__del__()
            jl.__del__()
        for o in (self.A_r, self.A_l, self.envop_jump, self.A, self.A_coh, self.A1, self.A2,
                  self.A3, self.envop, *self.env, *self.H):
            o.__del__()
        self.comp.close()
        self.ctx.close()


class RaggedSpinEnsemble:
    """Trajectories of driven spin chains of different lengths in one ragged pool.

    Trajectory i follows a chain of ``Ns[i]`` sites with positions
    ``models.positions(Ns[i], Ns[i], pos_seeds[i])`` (unit density, as
    ``SpinTrajectory`` with L = N) and draws r1, r2 from
    ``numpy.random.default_rng(seeds[i])``; all share the bond dimension D.
    The pool keeps the chains longest first (copy b holds trajectory
    ``order[b]``) in containers of max(Ns) slots (``Context.set_ragged``), so
    one batched RK4 call advances every chain: each GEMM, QR panel and
    environment update of a site slot is one launch for all chains present at
    that slot.  ``normalize_lo!`` of all chains is one call; the jump branch
    of a trajectory (spin_model_quantum_trajectory_rk4.jl:140-177) runs on a
    view of its copy, which is its own chain of Ns[i] sites.  A trajectory's
    arithmetic is that of the same trajectory alone in a pool of the same
    size, whatever the other chains' lengths.
    """

    def __init__(self, Ns, D, pos_seeds, seeds, dt=0.01, Omega=0.5, Gamma=1.0, k=2 * np.pi, gam_eg=0.0,
                 init=(1.0, 0.0), qr_rank_tol=0.0, precision="fp64", device=0):
        B = len(Ns)
        if not (len(pos_seeds) == len(seeds) == B and B >= 2):
            raise ValueError("one chain length, position seed and draw seed per trajectory, at least two")
        self.B, self.D, self.dt, self.gam_eg = B, D, dt, gam_eg
        self.Ns = [int(n) for n in Ns]
        self._model = (Omega, Gamma, k)
        self.order = sorted(range(B), key=lambda i: -self.Ns[i])
        self.copy_of = {i: b for b, i in enumerate(self.order)}
        ns = [self.Ns[i] for i in self.order]
        self.nmax = nmax = ns[0]
        dims = M._dims_rule(nmax, D, 2)
        for n in sorted(set(ns)):
            # the chain's own bonds through its slots, and one bond all across its gap
            h = n // 2
            if (M._dims_rule(n, D, 2) != [dims[i if i <= h else i + nmax - n] for i in range(n + 1)]
                    or len(set(dims[h:h + nmax - n + 1])) != 1):
                raise ValueError(f"a chain of {n} sites at D={D} is too short for a pool of {nmax}-site slots "
                                 "(its two parts would meet at an edge bond)")
        Hs = MD.spin_rk4_mpos(MD.positions(nmax, float(nmax), 0), 0.0, dt, Omega, Gamma, k)
        copy_bytes, ws, w = slab_sizes(dims, 2, Hs, 5, B, device)
        self.hbm_bytes = slab_bytes(copy_bytes, ws, B)
        self.ctx = M.Context.batched(B, copy_bytes, ws, device)
        self.ctx.set_ragged(ns)
        self.ctx.set_qr_rank_tol(qr_rank_tol)
        self.ctx.set_precision(precision)
        self.comp = self.ctx.companion()
        self.comp.set_qr_rank_tol(qr_rank_tol)
        self.H = [M.MPO(h, self.ctx) for h in Hs]
        self.A = M.MPS(nmax, 2, dims, self.ctx)
        self.A_coh, self.A1, self.A2, self.A3 = (self.A.similar() for _ in range(4))
        self.envop = M.build_env(dims, dims, [1] + [w] * (nmax - 1) + [1], ctx=self.ctx)
        self.env = [M.build_env(dims, dims, ctx=self.ctx) for _ in range(3)]
        self.Av = [self.A.view(b, self.comp) for b in range(B)]
        self.Acv = [self.A_coh.view(b, self.comp) for b in range(B)]
        self._init_vec = init
        self._per_len = {}       # chain length -> companion scratch of that length
        self._jumps = [None] * B  # per copy: (jr, jl) on the companion
        self.rngs = [None] * B
        self.p_coh = [[] for _ in range(B)]
        self.jumps = [[] for _ in range(B)]
        for b, i in enumerate(self.order):
            self._set_chain(b, pos_seeds[i])
            self.reset(i, seeds[i])
        self.t = 0.0

    def _scratch(self, n):
        s = self._per_len.get(n)
        if s is None:
            init, dims = M.mpsproductstate(n, self.D, 2, np.asarray(self._init_vec, dtype=complex), self.comp)
            s = {"init": init, "dims": dims,
                 "A_r": M.MPS(n, 2, dims, self.comp), "A_l": M.MPS(n, 2, dims, self.comp),
                 "envop_jump": M.build_env(dims, dims, [1] + [2] * (n - 1) + [1], ctx=self.comp)}
            self._per_len[n] = s
        return s

    def _set_chain(self, b, pos_seed):
        """Copy b's RK4 and jump MPOs for the chain with positions from pos_seed."""
        Omega, Gamma, k = self._model
        n = self.Ns[self.order[b]]
        z = MD.positions(n, float(n), pos_seed)
        for h_dev, h_host in zip(self.H, MD.spin_rk4_mpos(z, 0.0, self.dt, Omega, Gamma, k)):
            v = h_dev.view(b, self.comp)
            for j, site in enumerate(h_host):
                v.upload_site(j, site)
            v.__del__()
        old = self._jumps[b]
        jr, jl = MD.spin_jump_mpos(z, Gamma, k)
        self._jumps[b] = (M.MPO(jr, self.comp), M.MPO(jl, self.comp))
        if old is not None:
            for o in old:
                o.__del__()

    def reset(self, i, seed):
        """Restart trajectory i from the initial product state, with draws from default_rng(seed)."""
        b = self.copy_of[i]
        self.Av[b].copy_from(self._scratch(self.Ns[i])["init"])
        self.rngs[b] = np.random.default_rng(seed)
        self.p_coh[i] = []
        self.jumps[i] = []

    def launch(self):
        """Queue the one RK4 call that advances every chain (no host wait)."""
        M.RK4stp_apply_H_half_(self.dt, self.H, self.A_coh, self.A, self.A1, self.A2, self.A3, *self.env,
                               self.envop)

    def step(self):
        self.launch()
        return self.finish()

    def finish(self):
        """Batched ``normalize_lo!`` (waits for the stream), then every
        trajectory's draws and jump branch; returns p_coh by trajectory."""
        p = M.normalize_lo_batch(self.A_coh)
        out = np.empty(self.B)
        for b, i in enumerate(self.order):
            out[i] = p[b]
            self.p_coh[i].append(float(p[b]))
            r1, r2 = self.rngs[b].random(2)
            step = len(self.p_coh[i]) - 1
            if r1 < p[b]:
                continue
            n = self.Ns[i]
            s = self._scratch(n)
            A, Ac = self.Av[b], self.Acv[b]
            jr, jl = self._jumps[b]
            M.compress_var_apply_H_(s["A_r"], s["envop_jump"], A, A, jr, 1)
            p_r = M.normalize_lo_(s["A_r"])
            M.compress_var_apply_H_(s["A_l"], s["envop_jump"], A, A, jl, 1)
            p_l = M.normalize_lo_(s["A_l"])
            epop = M.measure_excitations_(None, A, HEE)
            acc = jump_table(epop, p_r, p_l, self.dt, self.gam_eg)
            kind, j = pick_jump(acc, r2, n)
            if kind == "eg":
                Ac.copy_from(A)
                M.apply_site_operator_(Ac, j, HGE)
                M.leftorthonormalizationQR_(Ac)
            elif kind == "r":
                Ac.copy_from(s["A_r"])
            else:
                Ac.copy_from(s["A_l"])
            self.jumps[i].append((step, kind, j))
        self.A.copy_from(self.A_coh)
        self.t += self.dt
        return out

    def state(self, i):
        """Trajectory i's sites (host copies)."""
        return self.Av[self.copy_of[i]].to_host()

    def sigma_z(self, i):
        """<sigma_z^j> of trajectory i."""
        return 2.0 * M.measure_excitations_(None, self.Av[self.copy_of[i]], HEE) - 1.0

    def close(self):
        for name in ("Av", "Acv"):
            for v in getattr(self, name):
                v.__del__()
        for s in self._per_len.values():
            for o in s.values():
                if hasattr(o, "__del__") and not isinstance(o, list):
                    o.__del__()
        for jj in self._jumps:
            if jj is not None:
                for o in jj:
                    o.__del__()
        for o in (self.A, self.A_coh, self.A1, self.A2, self.A3, self.envop, *self.env, *self.H):
            o.__del__()
        self.comp.close()
        self.ctx.close()
