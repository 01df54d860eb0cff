"""Writes tests/golden/rydberg_N100_D128_trace.npz: the FP64 CPU oracle's
trajectory of the benchmark workload (run from the repository root; about
one to two minutes per step on 8 host cores).

Workload: the Rydberg-ladder density matrix of rydberg_density_matrix_gpu.jl:
211-222 at N = 100, d = 9, 4 exponentials (MPO bond w = 14), D_max = 128,
dt = 0.005, positions uniform in [0, 100] from seed 1, started from
mpsgroundstate, RK4stp_apply_H_half! (10-argument form) with the QR rank
cutoff 1e-13, for ``STEPS`` steps.  Per step it records the trace
tr(rho) = scal_prod_no_conj(IDlmps, rho), the total populations
scal_prod_no_conj(N{E,S}lmps, rho) / tr, the Hilbert-Schmidt norm <rho|rho>
and three fingerprints <X_k|rho> against seeded vectorised product operators
X_k (expectation values of random product observables, which differ for any
two states that differ while the fixture stays small).

These are oracle outputs, not closed forms: the device is compared with them
to follow the oracle over many steps at the full size, where the oracle
itself is too slow to run inside a test.  The oracle is pinned separately
(closed forms, LAPACK, dense brute force; tests/test_oracle_*.py).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "matrixproductstates.jl_amd"))
from mpsamd import models  # noqa: E402
from oracle import mps_oracle as O  # noqa: E402

N, D, DT, N_EXP, SEED, TOL = 100, 128, 0.005, 4, 1, 1e-13
STEPS = int(os.environ.get("STEPS", "20"))


def fingerprint_mps(k, N=N):
    """Seeded vectorised product operator, site j = vec(1 + e_k H_j) with H_j a
    random Hermitian 3 x 3 matrix (e_k = 0.2, 0.5, 1.0): <X_k|rho> is then the
    expectation value of a product observable, of order tr(rho) and sensitive
    to every site's content."""
    rng = np.random.default_rng(1000 + k)
    eps = (0.2, 0.5, 1.0)[k]
    sites = []
    for _ in range(N):
        h = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
        a = np.eye(3) + eps * (h + h.conj().T) / 4
        sites.append(a.reshape(1, 9, 1, order="F").astype(np.complex128))
    return sites


def main():
    p = models.RydbergParams(n_exp=N_EXP)
    z = models.positions(N, float(N), SEED)
    L = models.rydberg_rk4_mpos(p, z, DT)
    O.set_qr_rank_tol(TOL)
    rho, dims = O.mpsgroundstate(N, D, 9)
    wop = [1] + [p.w] * (N - 1) + [1]
    envop = O.build_env_op(dims, dims, wop)
    e = [O.build_env(dims, dims) for _ in range(3)]
    r1, r2, r3 = ([a.copy() for a in rho] for _ in range(3))
    idm = models.identity_vec_mps(N)
    nem = models.sum_local_vec_mps(N, models.HEE3)
    nrm = models.sum_local_vec_mps(N, models.HSS3)
    X = [fingerprint_mps(k) for k in range(3)]
    rec = {k: [] for k in ("trace", "n_e", "n_r", "hs", "fp")}
    for s in range(STEPS):
        t0 = time.perf_counter()
        O.RK4stp_apply_H_half_(DT, L, rho, r1, r2, r3, *e, envop)
        tr = O.scal_prod_no_conj(idm, rho)
        rec["trace"].append(tr)
        rec["n_e"].append(O.scal_prod_no_conj(nem, rho) / tr)
        rec["n_r"].append(O.scal_prod_no_conj(nrm, rho) / tr)
        rec["hs"].append(O.scal_prod(rho, rho))
        rec["fp"].append([O.scal_prod(x, rho) for x in X])
        print(f"step {s + 1}: trace {tr:.15f} n_e {rec['n_e'][-1]:.15e} "
              f"({time.perf_counter() - t0:.1f} s)", flush=True)
        np.savez_compressed(os.path.join(HERE, "rydberg_N100_D128_trace.npz"),
                            steps=np.arange(1, s + 2), dt=DT, tol=TOL, v_max=p.v_max, u=p.u, kappa=p.kappa,
                            **{k: np.asarray(v, dtype=np.complex128) for k, v in rec.items()})


if __name__ == "__main__":
    main()
