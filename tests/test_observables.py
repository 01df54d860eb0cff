"""Photon-field observable builders (mpsamd.observables) against dense
operators, on the CPU.

* applyMPOtoMPO / conj_mpo / mpo_to_mps against dense products, adjoints and
  traces (MatrixProductStates.jl:67-99, 395-413; constructors.jl:298-314),
  and against the oracle's restatements of the same functions;
* I_r = scal_prod_no_conj(IRlmps, rho) equals tr(E_R^dag E_R rho) and I2_r
  equals tr(E_R^dag E_R^dag E_R E_R rho) for a random three-atom density
  matrix (rydberg_density_matrix_gpu.jl:135-151, 237-245), also with an input
  amplitude on the first site;
* ERProj (rydberg_density_matrix_gpu_g2.jl:117-126) is the superoperator
  rho -> E_R rho E_R^dag.
"""
import itertools

import numpy as np

from mpsamd import models as MD
from mpsamd import observables as OB
from oracle import dense as Dn
from oracle import mps_oracle as O


def _rand_mpo(rng, N, d, w):
    out = []
    for j in range(N):
        wl = 1 if j == 0 else w
        wr = 1 if j == N - 1 else w
        out.append(rng.standard_normal((wl, d, wr, d)) + 1j * rng.standard_normal((wl, d, wr, d)))
    return out


def test_mpo_algebra_against_dense(rng):
    a, b = _rand_mpo(rng, 3, 2, 2), _rand_mpo(rng, 3, 2, 3)
    A, B = Dn.mpo_to_dense(a), Dn.mpo_to_dense(b)
    assert np.allclose(Dn.mpo_to_dense(OB.applyMPOtoMPO(a, b)), B @ A, atol=1e-12)
    assert np.allclose(Dn.mpo_to_dense(OB.conj_mpo(a)), A.conj().T, atol=1e-12)
    # the oracle's restatements give the same tensors
    for x, y in zip(OB.applyMPOtoMPO(a, b), O.applyMPOtoMPO(a, b)):
        assert np.allclose(x, y, atol=1e-14)
    for x, y in zip(OB.mpo_to_mps(a), O.mpo_to_mps(a)):
        assert np.allclose(x, y, atol=0)


def _rand_density(rng, N):
    D = 3 ** N
    X = rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))
    rho = X @ X.conj().T
    return rho / np.trace(rho)


def _vec(rho, N):
    """The density-matrix MPS vector: site index s = bra + 3 ket (kron(A, id) acts
    as A rho), site 1 fastest; rho_vec[s] = rho[ket, bra]."""
    v = np.zeros(9 ** N, dtype=complex)
    for ket in itertools.product(range(3), repeat=N):
        for bra in itertools.product(range(3), repeat=N):
            s = sum((bra[j] + 3 * ket[j]) * 9 ** j for j in range(N))
            kg = sum(ket[j] * 3 ** j for j in range(N))
            bg = sum(bra[j] * 3 ** j for j in range(N))
            v[s] = rho[kg, bg]
    return v


def _field_dense(z, f_in, right=True):
    N = len(z)
    sgn = -1 if right else 1
    E = f_in * np.eye(3 ** N, dtype=complex)
    for j in range(N):
        E = E + 1j * np.sqrt(0.5) * np.exp(sgn * 1j * 2 * np.pi * z[j]) * Dn.site_op(N, 3, j, MD.HGE3)
    return E


def test_photon_intensities_are_traces(rng):
    N = 3
    z = MD.positions(N, 3.0, 2)
    rho = _rand_density(rng, N)
    rv = _vec(rho, N)
    for f_in in (0.0, 0.7):
        obs = OB.photon_observables(z, f_in=f_in)
        for name, right in (("R", True), ("L", False)):
            E = _field_dense(z, f_in if right else 0.0, right)
            n = E.conj().T @ E
            n2 = E.conj().T @ E.conj().T @ E @ E
            I1 = Dn.mps_to_dense(obs["I" + name]) @ rv
            I2 = Dn.mps_to_dense(obs["I2" + name]) @ rv
            assert abs(I1 - np.trace(n @ rho)) < 1e-12 * abs(np.trace(n @ rho))
            assert abs(I2 - np.trace(n2 @ rho)) < 1e-12 * abs(np.trace(n2 @ rho))
            # the same through the oracle's bilinear scalar product
            assert abs(O.scal_prod_no_conj(obs["I" + name], _mps_of(rv, N)) - I1) < 1e-12
        assert [s.shape[2] for s in obs["IR"][:-1]] == [4] * (N - 1)
        assert [s.shape[2] for s in obs["I2R"][:-1]] == [16] * (N - 1)


def _mps_of(v, N):
    """An exact MPS of a 9^N vector (site 1 fastest): the first site carries the
    whole vector, the others pass the remaining digits through."""
    out = [v.reshape(1, 9, 9 ** (N - 1), order="F")]
    Dl = 9 ** (N - 1)
    for _ in range(1, N):
        Dr = Dl // 9
        site = np.zeros((Dl, 9, Dr), dtype=complex)
        for a in range(Dl):
            site[a, a % 9, a // 9] = 1.0
        out.append(site)
        Dl = Dr
    return out


def test_er_projection_is_E_rho_Edag(rng):
    N = 3
    z = MD.positions(N, 3.0, 5)
    rho = _rand_density(rng, N)
    for f_amp in (0.0, 0.4):
        E = _field_dense(z, f_amp)
        P = Dn.mpo_to_dense(OB.er_projection_mpo(z, f_amp=f_amp))
        got = P @ _vec(rho, N)
        want = _vec(E @ rho @ E.conj().T, N)
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_g2():
    assert OB.g2(0.0, 1.0) is None
    assert OB.g2(2.0, 8.0) == 2.0
