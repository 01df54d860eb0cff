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
                 init=(1.0, 0.0), ctx: M.Context | None = None):
        self.N, self.D, self.dt, self.seed, self.gam_eg = N, D, dt, seed, gam_eg
        self.ctx = ctx if ctx is not None else M.default_context()
        self.z = MD.positions(N, float(N if L is None else L), seed)
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
