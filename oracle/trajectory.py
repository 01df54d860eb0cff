"""CPU oracle of the quantum-jump trajectory loop (spin_model_quantum_trajectory_rk4.jl:115-209).

TEST INFRASTRUCTURE ONLY: the checker for ``mpsamd.trajectory``.  Same branch
logic, same random draws (``numpy.random.default_rng(seed)``, r1 then r2 per
step), FP64 numpy/LAPACK arithmetic through ``oracle.mps_oracle``.
"""
import numpy as np

from oracle import mps_oracle as O

HGE = np.array([[0, 1], [0, 0]], dtype=complex)
HEE = np.array([[0, 0], [0, 1]], dtype=complex)


def measure_excitations(A, op=HEE):
    """spin_model_quantum_trajectory_rk4.jl:220-231 (right-to-left environment sweep)."""
    N = len(A)
    out = np.zeros(N)
    env = np.ones((1, 1), dtype=complex)
    for n in range(N - 1, -1, -1):
        Aenv = O.prod_LR_32(A[n], env)
        out[n] = O.local_exp(Aenv, A[n], op).real
        env = O.update_renv(A[n], Aenv)
    return out


class OracleSpinTrajectory:
    """Host restatement of one trajectory; inputs as for ``mpsamd.trajectory.SpinTrajectory``."""

    def __init__(self, N, D, seed, H, jr, jl, dt, gam_eg=0.0, init=(1.0, 0.0)):
        self.N, self.dt, self.gam_eg = N, dt, gam_eg
        self.H, self.jr, self.jl = H, jr, jl
        self.A, self.dims = O.mpsproductstate(N, D, 2, np.asarray(init, dtype=complex))
        self.envop = O.build_env_op(self.dims, self.dims, [1] + [4] * (N - 1) + [1])
        self.envop_jump = O.build_env_op(self.dims, self.dims, [1] + [2] * (N - 1) + [1])
        self.env = [O.build_env(self.dims, self.dims) for _ in range(3)]
        self.A1, self.A2, self.A3 = ([a.copy() for a in self.A] for _ in range(3))
        self.rng = np.random.default_rng(seed)
        self.p_coh, self.jumps = [], []

    def step(self):
        A_coh = [a.copy() for a in self.A]
        O.RK4stp_apply_H_half_(self.dt, self.H, self.A, self.A1, self.A2, self.A3, *self.env, self.envop,
                               psi0=A_coh)
        p = O.normalize_lo_(A_coh)
        self.p_coh.append(p)
        r1, r2 = self.rng.random(2)
        i = len(self.p_coh) - 1
        if r1 < p:
            self.A = A_coh
            return p
        A_r = [a.copy() for a in self.A]
        O.compress_var_apply_H_(A_r, self.envop_jump, self.A, self.A, self.jr, 1)
        p_r = O.normalize_lo_(A_r)
        A_l = [a.copy() for a in self.A]
        O.compress_var_apply_H_(A_l, self.envop_jump, self.A, self.A, self.jl, 1)
        p_l = O.normalize_lo_(A_l)
        epop = measure_excitations(self.A)
        N = self.N
        acc = np.zeros(N + 3)
        for j in range(N):
            acc[j + 1] = acc[j] + self.dt * self.gam_eg * epop[j]
        acc[N + 1] = acc[N] + self.dt * p_r
        acc[N + 2] = acc[N + 1] + self.dt * p_l
        acc = acc / acc[-1]
        if r2 < acc[N]:
            for j in range(N):
                if r2 < acc[j + 1]:
                    O.apply_site_operator_(self.A[j], HGE)
                    O.leftorthonormalizationQR_(self.A)
                    self.jumps.append((i, "eg", j))
                    break
        elif r2 < acc[N + 1]:
            self.A = A_r
            self.jumps.append((i, "r", -1))
        else:
            self.A = A_l
            self.jumps.append((i, "l", -1))
        return p

    def sigma_z(self):
        return 2.0 * measure_excitations(self.A) - 1.0
