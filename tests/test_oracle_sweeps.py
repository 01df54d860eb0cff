"""Pin the oracle's sweeps and RK4 against dense linear algebra.

At full bond dimension the one-site projections are exact, so every sweep
must reproduce the dense result; RK4 must reproduce dense RK4 and, for the
drive-free chain, the closed-form single-excitation recurrence.
"""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle import mps_oracle as O
from oracle import dense as Dn
from mpsamd import models
from helpers import crandn, random_mpo, copy_sites, rel


def full_random(N, d, seed):
    return O.mpsrandom(N, d ** N, d, np.random.default_rng(seed))


def test_compress_var_apply_H_left_exact(rng):
    N, d = 6, 2
    psi, dims = full_random(N, d, 1)
    H = random_mpo(rng, N, d, 3)
    A0 = copy_sites(psi)
    O.compress_var_apply_H_left_(A0, O.build_env_op(dims, dims, [1] + [3] * (N - 1) + [1]), psi, psi, H)
    assert rel(Dn.mps_to_dense(A0), Dn.mpo_to_dense(H) @ Dn.mps_to_dense(psi)) < 1e-12


@pytest.mark.parametrize("numv", [2, 4])
def test_compress_sum_var_apply_H_exact(numv, rng):
    N, d = 5, 2
    states = [full_random(N, d, 10 + l)[0] for l in range(numv)]
    Ain, dims = full_random(N, d, 3)
    H = random_mpo(rng, N, d, 2)
    coef = list(crandn(rng, numv))
    target = sum(coef[l] * Dn.mps_to_dense(states[l]) for l in range(numv - 1))
    target = target + coef[-1] * (Dn.mpo_to_dense(H) @ Dn.mps_to_dense(states[-1]))
    wd = [1] + [2] * (N - 1) + [1]
    A0 = copy_sites(Ain)
    O.compress_sum_var_apply_H_left_(A0, [O.build_env(dims, dims) for _ in range(numv - 1)],
                                    O.build_env_op(dims, dims, wd), Ain, coef, states, H)
    assert rel(Dn.mps_to_dense(A0), target) < 1e-12
    # _right!: the H term enters through the conjugated left env -> conj(coef)
    Ainr = copy_sites(Ain)
    O.rightorthonormalizationQR_(Ainr)
    A0 = copy_sites(Ainr)
    O.compress_sum_var_apply_H_right_(A0, [O.build_env(dims, dims) for _ in range(numv - 1)],
                                     O.build_env_op(dims, dims, wd), Ainr, coef, states, H)
    target_r = target - coef[-1] * (Dn.mpo_to_dense(H) @ Dn.mps_to_dense(states[-1])) \
        + np.conj(coef[-1]) * (Dn.mpo_to_dense(H) @ Dn.mps_to_dense(states[-1]))
    assert rel(Dn.mps_to_dense(A0), target_r) < 1e-12


def test_compress_var_and_sum_var_exact(rng):
    N, d = 5, 2
    t, dims = full_random(N, d, 4)
    g, _ = full_random(N, d, 5)
    A0 = copy_sites(g)
    O.compress_var_(A0, t, O.build_env(dims, dims), 1)
    assert rel(Dn.mps_to_dense(A0), Dn.mps_to_dense(t)) < 1e-12
    vecs = [full_random(N, d, 20 + i)[0] for i in range(3)]
    coef = [0.5, -1.0j, 0.25]
    out = O.compress_sum_var(g, coef, vecs, 1)
    ref = sum(c * Dn.mps_to_dense(v) for c, v in zip(coef, vecs))
    assert rel(Dn.mps_to_dense(out), ref) < 1e-12


def test_compress_var_apply_H_and_sum_var_apply_H_exact(rng):
    N, d = 5, 2
    psi, dims = full_random(N, d, 6)
    H = random_mpo(rng, N, d, 2)
    wd = [1] + [2] * (N - 1) + [1]
    A0 = copy_sites(psi)
    O.compress_var_apply_H_(A0, O.build_env_op(dims, dims, wd), psi, psi, H, 1)
    assert rel(Dn.mps_to_dense(A0), Dn.mpo_to_dense(H) @ Dn.mps_to_dense(psi)) < 1e-12
    vecs = [full_random(N, d, 30 + i)[0] for i in range(2)]
    coef = [0.7, 0.2 - 0.1j]
    A0 = copy_sites(psi)
    O.compress_sum_var_apply_H_(A0, O.build_env_op(dims, dims, wd), coef, vecs, H, 1)
    ref = coef[0] * Dn.mps_to_dense(vecs[0]) + coef[1] * Dn.mpo_to_dense(H) @ Dn.mps_to_dense(vecs[1])
    assert rel(Dn.mps_to_dense(A0), ref) < 1e-12


def _rk4_mps(H, psi, dims, dt, steps, wop):
    psi = copy_sites(psi)
    envop = O.build_env_op(dims, dims, wop)
    e = [O.build_env(dims, dims) for _ in range(3)]
    p1, p2, p3 = copy_sites(psi), copy_sites(psi), copy_sites(psi)
    for _ in range(steps):
        O.RK4stp_apply_H_half_(dt, H, psi, p1, p2, p3, *e, envop)
    return psi


def test_rk4_equals_dense_rk4():
    N, dt, steps = 6, 0.02, 5
    z = models.positions(N, 6.0, 2)
    H = models.spin_rk4_mpos(z, 0.0, dt, Omega=0.7)
    psi, dims = O.mpsproductstate(N, 2 ** N, 2, np.array([1, 0], dtype=complex))
    got = Dn.mps_to_dense(_rk4_mps(H, psi, dims, dt, steps, [1] + [4] * (N - 1) + [1]))
    Mx = Dn.mpo_to_dense(models.spin_mpo(z, 0.7, 1.0, 2 * np.pi, con=0.0, delt=1.0))
    v = Dn.mps_to_dense(psi)
    for _ in range(steps):
        v = Dn.rk4_dense(Mx, v, dt)
    assert rel(got, v) < 1e-12


def test_rk4_single_excitation_closed_form():
    # Omega = 0: one excitation stays in the N-dim sector, c <- P4(-i dt M) c with
    # M_ij = -i Gamma/2 exp(i k |z_i - z_j|) (the effective Hamiltonian's matrix).
    N, dt, steps = 6, 0.05, 8
    z = models.positions(N, 6.0, 7)
    H = models.spin_rk4_mpos(z, 0.0, dt, Omega=0.0)
    st = lambda n: np.array([0, 1], dtype=complex) if n == 1 else np.array([1, 0], dtype=complex)
    psi, dims = O.mpsproductstate(N, 2 ** N, 2, st)
    got = Dn.mps_to_dense(_rk4_mps(H, psi, dims, dt, steps, [1] + [4] * (N - 1) + [1]))
    Heff = models.spin_heff_dense_matrix(z, 0.0, 1.0, 2 * np.pi)
    G = -1j * Heff
    c = np.zeros(N, dtype=complex)
    c[0] = 1.0
    P = np.eye(N)
    term = np.eye(N, dtype=complex)
    for k in range(1, 5):
        term = term @ (dt * G) / k
        P = P + term
    for _ in range(steps):
        c = P @ c
    # dense index of |e at site j> is 2^j (site 1 fastest)
    amp = np.array([got[2 ** j] for j in range(N)])
    assert np.abs(amp - c).max() < 1e-12
    mask = np.ones(2 ** N, bool)
    mask[[2 ** j for j in range(N)]] = False
    assert np.abs(got[mask]).max() < 1e-13


def test_rydberg_lindblad_rk4_trace_and_dense():
    N, dt, steps = 3, 0.01, 4
    p = models.RydbergParams(n_exp=2)
    z = models.positions(N, 3.0, 1)
    L = models.rydberg_rk4_mpos(p, z, dt)
    rho, dims = O.mpsgroundstate(N, 81, 9)
    got = _rk4_mps(L, rho, dims, dt, steps, [1] + [p.w] * (N - 1) + [1])
    I = models.identity_vec_mps(N)
    assert abs(O.scal_prod_no_conj(I, got) - 1.0) < 1e-12
    Ld = Dn.mpo_to_dense(models.rydberg_lindblad_mpo(p, z, 0.0, 1.0))
    v = Dn.mps_to_dense(rho)
    for _ in range(steps):
        v = Dn.rk4_dense(Ld, v, dt)
    assert rel(Dn.mps_to_dense(got), v) < 1e-12


def test_arnoldi_expm_matches_expm(rng):
    N, d, numv = 4, 2, 6
    psi, dims = O.mpsrandom(N, 16, d, rng)
    psi = [s.real.astype(complex) for s in psi]
    O.leftorthonormalizationQR_(psi)
    H = [np.real(h).astype(complex) * 0.3 for h in random_mpo(rng, N, d, 2)]
    rhov = [copy_sites(psi) for _ in range(numv)]
    O.Arnoldi_expm_(H, rhov, numv, None, O.build_env_op(dims, dims, [1, 2, 2, 2, 1]))
    ref = expm(Dn.mpo_to_dense(H)) @ Dn.mps_to_dense(psi)
    ref = ref / np.linalg.norm(ref)
    got = Dn.mps_to_dense(rhov[0])
    assert rel(got, ref) < 1e-6


def test_compression_svd_schmidt(rng):
    N, d = 6, 2
    psi, dims = full_random(N, d, 9)
    O.leftorthonormalizationQR_(psi)
    v = Dn.mps_to_dense(psi)
    work = copy_sites(psi)
    err_glob, err_trim, ent = O.compressionSVD_(work, 64, 1e-15)
    assert abs(err_glob) < 1e-12
    # entanglement of the cut left of site 3 from the dense Schmidt values
    s = np.linalg.svd(v.reshape(2 ** 3, 2 ** 3, order="F"), compute_uv=False)
    assert abs(ent[3] - (-2 * np.sum(s ** 2 * np.log(s)))) < 1e-10
    assert rel(Dn.mps_to_dense(work), v) < 1e-12
    # truncation to dmax = 2: kept weight = sum of the 2 largest s^2 at each cut
    work = copy_sites(psi)
    eg, et, _ = O.compressionSVD_(work, 2, 1e-15)
    assert max(w.shape[0] for w in work) <= 2
    assert 0 < eg < 1


def test_scalar_products_dense(rng):
    N, d = 5, 3
    a, _ = O.mpsrandom(N, 9, d, rng)
    b, _ = O.mpsrandom(N, 9, d, rng)
    va, vb = Dn.mps_to_dense(a), Dn.mps_to_dense(b)
    H = random_mpo(rng, N, d, 2)
    assert np.isclose(O.scal_prod(a, b), np.vdot(va, vb), atol=1e-13)
    assert np.isclose(O.scal_prod_no_conj(a, b), np.sum(va * vb), atol=1e-13)
    assert np.isclose(O.scal_op_prod(a, H, b), np.conj(np.vdot(va, Dn.mpo_to_dense(H) @ vb)), atol=1e-13)
    assert np.isclose(O.normMPS(a), 1.0)
    assert abs(O.diff_norm(a, a)) < 1e-12


def test_measure_excitations_dense(rng):
    N, d = 5, 2
    a, _ = O.mpsrandom(N, 8, d, rng)
    v = Dn.mps_to_dense(a)
    env = np.ones((1, 1), dtype=complex)
    for n in range(N - 1, -1, -1):
        Aenv = O.prod_LR_32(a[n], env)
        e = O.local_exp(Aenv, a[n], models.HEE2).real
        env = O.update_renv(a[n], Aenv)
        assert abs(e - np.vdot(v, Dn.site_op(N, 2, n, models.HEE2) @ v).real) < 1e-12
