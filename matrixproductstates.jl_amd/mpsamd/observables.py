"""Photon-field observables of the density-matrix workloads (host-side builders).

The reference measures the light a Rydberg-ladder chain emits into the
waveguide with vectorised operator MPS (``examples/rydberg_density_matrix_gpu.jl:
135-151, 237-245``):

* ``ERmpo`` / ``ELmpo`` -- bond-2 MPOs of the right- and left-moving field
  operators E_R = f + sum_j i sqrt(Gamma_j/2) exp(-i k z_j) sigma_j^-
  (E_L with exp(+i k z_j) and no input term), built with ``makempo``;
* ``IRmpo = applyMPOtoMPO(ERmpo, conj_mpo(ERmpo))`` = E_R^dag E_R and
  ``IR2mpo = applyMPOtoMPO(applyMPOtoMPO(ERmpo, IRmpo), conj_mpo(ERmpo))`` =
  E_R^dag E_R^dag E_R E_R (likewise for E_L);
* ``mpo_to_mps`` turns each into a d^2 = 9 MPS, and
  ``scal_prod_no_conj(IRlmps, rho)`` is the intensity I_r = tr(E_R^dag E_R rho)
  (I2_r, I_l, I2_l the same way); the second-order correlation of the
  right-moving light is g2 = I2_r / I_r^2.

``rydberg_density_matrix_gpu_g2.jl:117-131, 213`` also builds ``ERProj``, the
bond-4 superoperator MPO of rho -> E_R rho E_R^dag, and conditions the state on
a detected photon with ``compress_var_apply_H!(rho, envproj, rho, rho1, ERProj, 4)``.

Everything here is a host (numpy) construction of small MPOs and MPS in the
reference's layouts (MPO ``[wl, s_out, wr, s_in]``, MPS ``[Dl, s, Dr]``,
column-major); the measurements themselves are device scalar products.
"""
from __future__ import annotations

import numpy as np

from . import models as MD

CDT = np.complex128


def makempo(mpomat, N):
    """makempo (constructors.jl:229-253): a b1 x b2 matrix of per-site operator
    functions f(n) (1-based site) -> MPO sites [b1, d, b2, d]; the first site
    keeps row 1, the last site column b2."""
    b1, b2 = len(mpomat), len(mpomat[0])
    sites = []
    for n in range(1, N + 1):
        d = np.asarray(mpomat[0][0](n)).shape[0]
        s = np.zeros((b1, d, b2, d), dtype=CDT)
        for l in range(b1):
            for m in range(b2):
                s[l, :, m, :] = mpomat[l][m](n)
        sites.append(s)
    sites[0] = sites[0][0:1].copy()
    sites[-1] = sites[-1][:, :, b2 - 1:b2, :].copy()
    return sites


def conj_site_mpo(m):
    """conj_site_mpo (MatrixProductStates.jl:410-413): the adjoint of one site,
    conj and swap of the two physical indices."""
    return np.ascontiguousarray(np.conj(np.transpose(m, (0, 3, 2, 1))))


def conj_mpo(mpo):
    """conj_mpo (MatrixProductStates.jl:395-407): the MPO of the adjoint."""
    return [conj_site_mpo(m) for m in mpo]


def apply_site_MPOtoMPO(mpo1, mpo2):
    """apply_site_MPOtoMPO (MatrixProductStates.jl:91-99): one site of the
    operator product mpo2 * mpo1, bonds combined with mpo1's index fastest:
    out[a1 + wl1 a2, s, b1 + wr1 b2, t] = sum_u mpo2[a2, s, b2, u] mpo1[a1, u, b1, t]."""
    wl1, d, wr1, _ = mpo1.shape
    wl2, _, wr2, _ = mpo2.shape
    t = np.einsum("ASBu,aubt->aASbBt", mpo2, mpo1)
    return t.reshape(wl1 * wl2, d, wr1 * wr2, d, order="F")


def applyMPOtoMPO(mpo1, mpo2):
    """applyMPOtoMPO (MatrixProductStates.jl:67-76): the MPO of mpo2 * mpo1."""
    return [apply_site_MPOtoMPO(a, b) for a, b in zip(mpo1, mpo2)]


def mpo_to_mps_site(m):
    """mpo_to_mps_site (constructors.jl:309-314): [wl, s, wr, t] -> [wl, s + d t, wr]."""
    wl, d, wr, _ = m.shape
    return np.ascontiguousarray(np.transpose(m, (0, 1, 3, 2))).reshape(wl, d * d, wr, order="F")


def mpo_to_mps(mpo):
    """mpo_to_mps (constructors.jl:298-307)."""
    return [mpo_to_mps_site(m) for m in mpo]


def field_mpo(z, Gamma_1d=1.0, k=2 * np.pi, right=True, f_in=0.0):
    """ERmpo (right=True) / ELmpo of rydberg_density_matrix_gpu.jl:135-146 for
    the three-level ladder (levels g, e, r): bond 2,
    E = f_in + sum_j i sqrt(Gamma_1d/2) exp(-+ i k z_j) |g><e|_j, with the input
    amplitude f_in on the first site (the reference writes f(t) there for E_R)."""
    z = np.asarray(z, dtype=np.float64)
    N = len(z)
    sgn = -1.0 if right else 1.0
    c = 1j * np.sqrt(Gamma_1d / 2) * np.exp(sgn * 1j * k * z)
    one = lambda n: MD.ID3   # noqa: E731
    zero = lambda n: 0 * MD.ID3   # noqa: E731
    mpo = makempo([[one, lambda n: c[n - 1] * MD.HGE3], [zero, one]], N)
    if f_in != 0.0:
        mpo[0][0, :, 1, :] = f_in * MD.ID3 + c[0] * MD.HGE3
    return mpo


def photon_observables(z, Gamma_1d=1.0, k=2 * np.pi, f_in=0.0):
    """The vectorised operator MPS IRlmps, IR2lmps, ILlmps, IL2lmps
    (rydberg_density_matrix_gpu.jl:135-151): bond dimensions 4, 16, 4, 16."""
    ER = field_mpo(z, Gamma_1d, k, True, f_in)
    EL = field_mpo(z, Gamma_1d, k, False)
    out = {}
    for name, E in (("R", ER), ("L", EL)):
        I1 = applyMPOtoMPO(E, conj_mpo(E))
        I2 = applyMPOtoMPO(applyMPOtoMPO(E, I1), conj_mpo(E))
        out["I" + name] = mpo_to_mps(I1)
        out["I2" + name] = mpo_to_mps(I2)
    return out


def g2(I, I2):
    """Second-order correlation of the emitted light, I2 / I^2 (None if I = 0)."""
    return None if I == 0 else I2 / I ** 2


def er_projection_mpo(z, Gamma_1d=1.0, k=2 * np.pi, f_amp=0.0):
    """ERProj (rydberg_density_matrix_gpu_g2.jl:117-126): the bond-4 MPO on the
    vectorised site index (d^2 = 9) of the superoperator rho -> E_R rho E_R^dag,
    E_R = f_amp + sum_j c_j |g><e|_j, c_j = i sqrt(Gamma_1d/2) exp(-i k z_j):
    kron(A, id) acts as A rho, kron(id, B) as rho B^T."""
    z = np.asarray(z, dtype=np.float64)
    N = len(z)
    c = 1j * np.sqrt(Gamma_1d / 2) * np.exp(-1j * k * z)
    I9 = np.kron(MD.ID3, MD.ID3)
    ket = np.kron(MD.HGE3, MD.ID3)    # sigma^- rho
    bra = np.kron(MD.ID3, MD.HGE3)    # rho sigma^+
    both = np.kron(MD.HGE3, MD.HGE3)  # sigma^- rho sigma^+
    Z = 0 * I9
    m = [[lambda n: I9, lambda n: c[n - 1] * ket, lambda n: np.conj(c[n - 1]) * bra,
          lambda n: abs(c[n - 1]) ** 2 * both],
         [lambda n: Z, lambda n: I9, lambda n: Z, lambda n: np.conj(c[n - 1]) * bra],
         [lambda n: Z, lambda n: Z, lambda n: I9, lambda n: c[n - 1] * ket],
         [lambda n: Z, lambda n: Z, lambda n: Z, lambda n: I9]]
    mpo = makempo(m, N)
    if f_amp != 0.0:
        f = complex(f_amp)
        mpo[0][0, :, 1, :] = f * I9 + c[0] * ket
        mpo[0][0, :, 2, :] = np.conj(f) * I9 + np.conj(c[0]) * bra
        mpo[0][0, :, 3, :] = (abs(c[0]) ** 2 * both + abs(f) ** 2 * I9 + np.conj(f) * c[0] * ket
                              + f * np.conj(c[0]) * bra)
    return mpo
