"""Host-side model builders: the MPOs and initial states the hot path consumes.

These restate the example drivers of jdnz/MatrixProductStates.jl (they are
inputs to the path, not part of it):

* ``spin_mpo`` -- the exact bond-4 MPO of ``con*1 + delt*(-i H_eff)`` for a
  driven two-level chain coupled to a waveguide, following
  ``examples/spin_model_quantum_trajectory_rk4.jl:235-293`` (``sm_hamiltonian``
  / ``local_h``), with random atom positions and a coherent drive Omega.
* ``rydberg_lindblad_mpo`` -- the vectorised Lindbladian MPO of a
  three-level ladder (g, e, r) with collective waveguide decay and a Rydberg
  interaction fitted by a sum of exponentials, following
  ``examples/rydberg_density_matrix_gpu.jl:318-420`` (``construct_L_Ryd`` /
  ``local_l``).  The reference fits the interaction in site-index distance
  (constant ``lam`` per site); here atoms sit at random positions, so every
  bond carries its own factor ``exp(-kappa_e * (z[j+1] - z[j]))``.

All arrays are complex128 in the reference layout ``O[wl, s_out, wr, s_in]``.
The vectorisation convention is the reference's: ``kron(A, id)`` acts as
``A rho`` and ``kron(id, B)`` as ``rho B^T`` on a site's 9-dimensional index.
"""
from __future__ import annotations

import numpy as np

CDT = np.complex128

# two-level operators (spin_model_quantum_trajectory_rk4.jl:34-38)
HGG2 = np.array([[1, 0], [0, 0]], dtype=CDT)
HEE2 = np.array([[0, 0], [0, 1]], dtype=CDT)
HGE2 = np.array([[0, 1], [0, 0]], dtype=CDT)   # |g><e| = sigma^-
HEG2 = HGE2.T.copy()                            # sigma^+
ID2 = np.eye(2, dtype=CDT)


def _op3(i, j):
    m = np.zeros((3, 3), dtype=CDT)
    m[i, j] = 1.0
    return m


# three-level operators, levels (g, e, r) (rydberg_density_matrix_gpu.jl:114-121)
HGG3, HEE3, HSS3 = _op3(0, 0), _op3(1, 1), _op3(2, 2)
HGE3, HEG3 = _op3(0, 1), _op3(1, 0)
HES3, HSE3 = _op3(1, 2), _op3(2, 1)
ID3 = np.eye(3, dtype=CDT)
ID9 = np.eye(9, dtype=CDT)


def positions(N, L, seed):
    """Sorted uniform atom positions in [0, L] (numpy PCG64 stream, ``seed``)."""
    rng = np.random.default_rng(seed)
    return np.sort(rng.uniform(0.0, L, N))


# --------------------------------------------------------------------------
# driven two-level chain (trajectory model)
# --------------------------------------------------------------------------

def spin_local_h(N, Omega, Gamma, con, delt):
    """``local_h`` (spin_model_quantum_trajectory_rk4.jl:287-293) with a coherent drive.

    con/N * 1 + delt * (-i H_loc) with H_loc = -Omega/2 sigma_x - i Gamma/2 |e><e|.
    """
    return con * ID2 / N + delt * (-(Gamma / 2) * HEE2 + 1j * (Omega / 2) * (HEG2 + HGE2))


def spin_mpo(z, Omega=0.5, Gamma=1.0, k=2 * np.pi, con=1.0, delt=0.01):
    """Bond-4 MPO of con*1 + delt*(-i H_eff) (``sm_hamiltonian``, :235-271).

    H_eff = -Omega/2 sum sigma_x - i Gamma/2 sum_ij exp(i k |z_i - z_j|) sigma_i^+ sigma_j^-.
    """
    z = np.asarray(z, dtype=np.float64)
    N = len(z)
    ph = np.exp(1j * k * np.diff(z))
    cp = np.sqrt(delt * Gamma / 2) * np.ones(N)
    loc = spin_local_h(N, Omega, Gamma, con, delt)
    H = []
    for j in range(N):
        if j == 0:
            W = np.zeros((1, 2, 4, 2), dtype=CDT)
            W[0, :, 0, :] = ID2
            W[0, :, 1, :] = -cp[j] * ph[j] * HEG2
            W[0, :, 2, :] = -cp[j] * ph[j] * HGE2
            W[0, :, 3, :] = loc
        elif j == N - 1:
            W = np.zeros((4, 2, 1, 2), dtype=CDT)
            W[0, :, 0, :] = loc
            W[1, :, 0, :] = cp[j] * HGE2
            W[2, :, 0, :] = cp[j] * HEG2
            W[3, :, 0, :] = ID2
        else:
            W = np.zeros((4, 2, 4, 2), dtype=CDT)
            W[0, :, 0, :] = ID2
            W[0, :, 1, :] = -cp[j] * ph[j] * HEG2
            W[0, :, 2, :] = -cp[j] * ph[j] * HGE2
            W[0, :, 3, :] = loc
            W[1, :, 3, :] = cp[j] * HGE2
            W[2, :, 3, :] = cp[j] * HEG2
            W[1, :, 1, :] = ph[j] * ID2
            W[2, :, 2, :] = ph[j] * ID2
            W[3, :, 3, :] = ID2
        H.append(W)
    return H


def spin_heff_dense_matrix(z, Omega, Gamma, k):
    """Single-excitation-sector matrix and the full dense H_eff (for small N checks)."""
    z = np.asarray(z, dtype=np.float64)
    N = len(z)
    M = -0.5j * Gamma * np.exp(1j * k * np.abs(z[:, None] - z[None, :]))
    return M


def spin_rk4_mpos(z, t, dt, Omega=0.5, Gamma=1.0, k=2 * np.pi):
    """[H1, H2, H3] of the RK4 half-step (spin_model_quantum_trajectory_rk4.jl:110-121).

    H1 = H3 = 1 + (dt/2)(-i H_eff), H2 = -i H_eff (time-independent drive, so t is unused).
    """
    H1 = spin_mpo(z, Omega, Gamma, k, con=1.0, delt=dt / 2)
    H2 = spin_mpo(z, Omega, Gamma, k, con=0.0, delt=1.0)
    H3 = spin_mpo(z, Omega, Gamma, k, con=1.0, delt=dt / 2)
    return [H1, H2, H3]


def spin_jump_mpos(z, Gamma=1.0, k=2 * np.pi):
    """Bond-2 right/left waveguide jump MPOs (spin_model_quantum_trajectory_rk4.jl:51-55).

    E_R = sum_j i sqrt(Gamma/2) exp(-i k z_j) sigma_j^-, E_L with exp(+i k z_j).
    """
    z = np.asarray(z, dtype=np.float64)
    N = len(z)
    out = []
    for sgn in (-1, +1):
        mpo = []
        for j in range(N):
            c = 1j * np.sqrt(Gamma / 2) * np.exp(sgn * 1j * k * z[j])
            W = np.zeros((2, 2, 2, 2), dtype=CDT)
            W[0, :, 0, :] = ID2
            W[0, :, 1, :] = c * HGE2
            W[1, :, 1, :] = ID2
            if j == 0:
                W = W[0:1].copy()
            if j == N - 1:
                W = W[:, :, 1:2, :].copy()
            mpo.append(W)
        out.append(mpo)
    return out[0], out[1]


# --------------------------------------------------------------------------
# three-level Rydberg ladder, vectorised Lindbladian (density-matrix model)
# --------------------------------------------------------------------------

def fit_c6_exponentials(n_exp=4, C6=1.0, r_lo=1.0, r_hi=10.0, npts=400, v_max=200.0):
    """Fit C6/r^6 on [r_lo, r_hi] by sum_e u_e exp(-kappa_e r) (relative least squares).

    Deterministic: fixed log-spaced start, scipy ``least_squares`` on
    (log u, log kappa).  Below ``r_lo`` the fitted potential saturates at
    sum(u) (a soft core).  One more residual row pins that core to ``v_max``
    (None: unpinned).  The core sets the fastest frequencies of the
    Lindbladian: a cluster of close atoms, all in the Rydberg level, has ket-bra
    coherences that rotate at the sum of its pair energies, and RK4 amplifies a
    mode whose dt*lambda leaves its stability region.  On the N=100 chain at
    positions uniform in [0, 100] (closest pair 0.018 apart) and dt = 0.005,
    the unpinned fit's core of 469 puts three-atom windows at
    max|P4(dt lambda)| = 1.72 and the FP64 evolution diverges within 40 steps;
    a core of 200 keeps every three-atom window, and the tightest four-atom one,
    at max|P4| = 1 (tail error 9% on [r_lo, r_hi] against 2.5% unpinned).
    Returns (u, kappa, max relative tail error).
    """
    from scipy.optimize import least_squares

    r = np.geomspace(r_lo, r_hi, npts)
    target = C6 / r ** 6
    kap0 = np.geomspace(6.0 / r_hi, 6.0 / r_lo, n_exp)
    A = np.exp(-np.outer(r, kap0))
    u0 = np.linalg.lstsq(A / target[:, None], np.ones_like(r), rcond=None)[0]
    u0 = np.abs(u0) + 1e-6

    def tail(p):
        u = np.exp(p[:n_exp])
        kap = np.exp(p[n_exp:])
        return (np.exp(-np.outer(r, kap)) @ u) / target - 1.0

    def res(p):
        if v_max is None:
            return tail(p)
        return np.concatenate([tail(p), [20.0 * (np.exp(p[:n_exp]).sum() / v_max - 1.0)]])

    sol = least_squares(res, np.concatenate([np.log(u0), np.log(kap0)]), method="lm",
                        xtol=1e-14, ftol=1e-14, max_nfev=20000)
    u = np.exp(sol.x[:n_exp])
    kap = np.exp(sol.x[n_exp:])
    order = np.argsort(kap)
    return u[order], kap[order], float(np.max(np.abs(tail(sol.x))))


def _sl(A):
    """Superoperator of rho -> A rho on one site (reference: kron(A, id))."""
    return np.kron(A, ID3)


def _sr(B):
    """Superoperator of rho -> rho B on one site (reference: kron(id, B^T))."""
    return np.kron(ID3, B.T)


class RydbergParams:
    """Parameters of the three-level ladder chain (defaults: the N=100, d=9 density-matrix workload)."""

    def __init__(self, Omega_p=0.2, Omega_c=1.0, gamma_e=1.0, gamma_r=0.01, Gamma_1d=1.0,
                 k=2 * np.pi, C6=1.0, n_exp=4, fit_range=(1.0, 10.0), v_max=200.0):
        self.Omega_p = Omega_p
        self.Omega_c = Omega_c
        self.gamma_e = gamma_e
        self.gamma_r = gamma_r
        self.Gamma_1d = Gamma_1d
        self.k = k
        self.C6 = C6
        self.n_exp = n_exp
        self.v_max = v_max
        self.u, self.kappa, self.fit_err = fit_c6_exponentials(n_exp, C6, *fit_range, v_max=v_max)

    @property
    def w(self):
        return 6 + 2 * self.n_exp


def rydberg_local_l(p: RydbergParams, N, con, delt):
    """``local_l`` (rydberg_density_matrix_gpu.jl:405-420) for the ladder parameters ``p``.

    con/N * 1 + delt * (L_loc), L_loc = -i H_eff rho + i rho H_eff^dag
    + (gamma_e + Gamma_1d) s^- rho s^+ + gamma_r P_r rho P_r.
    """
    Hloc = -(p.Omega_p / 2) * (HEG3 + HGE3) - (p.Omega_c / 2) * (HSE3 + HES3)
    Heff = Hloc - 0.5j * (p.gamma_e + p.Gamma_1d) * HEE3 - 0.5j * p.gamma_r * HSS3
    L = (_sl(-1j * Heff) + _sr(1j * Heff.conj().T)
         + (p.gamma_e + p.Gamma_1d) * np.kron(HGE3, HGE3) + p.gamma_r * np.kron(HSS3, HSS3))
    return con / N * ID9 + delt * L


def rydberg_lindblad_mpo(p: RydbergParams, z, con, delt):
    """Vectorised Lindbladian MPO of con*1 + delt*L (``construct_L_Ryd``, :318-391).

    Bond layout (w = 6 + 2 n_exp): 0 identity/start, 1..4 the four waveguide
    channels, 5.. the ket/bra Rydberg channels per exponential, w-1 the end.
    """
    z = np.asarray(z, dtype=np.float64)
    N = len(z)
    w = p.w
    dz = np.diff(z)
    ph = np.exp(1j * p.k * dz)
    cp = np.sqrt(delt * p.Gamma_1d / 2) * np.ones(N)
    cpc = np.conj(cp)
    lam = np.exp(-np.outer(dz, p.kappa))            # [bond j, e]: factor for bond j -> j+1
    sk = np.sqrt(-1j * delt * p.u + 0j)              # ket-side Rydberg coefficients
    sb = np.sqrt(1j * delt * p.u + 0j)               # bra-side
    ket_ss, bra_ss = _sl(HSS3), _sr(HSS3)
    I = ID9
    last = w - 1
    mpo = []
    for j in range(N):
        W = np.zeros((w, 9, w, 9), dtype=CDT)
        loc = rydberg_local_l(p, N, con, delt)
        if j < N - 1:                                # start a coupling on this site
            W[0, :, 0, :] = I
            W[0, :, 1, :] = cp[j] * ph[j] * (_sr(HEG3) - _sl(HEG3))
            W[0, :, 2, :] = cp[j] * ph[j] * _sl(HGE3)
            W[0, :, 3, :] = cpc[j] * np.conj(ph[j]) * _sr(HEG3)
            W[0, :, 4, :] = cpc[j] * np.conj(ph[j]) * (_sl(HGE3) - _sr(HGE3))
            for e in range(p.n_exp):
                W[0, :, 5 + 2 * e, :] = sk[e] * lam[j, e] * ket_ss
                W[0, :, 6 + 2 * e, :] = sb[e] * lam[j, e] * bra_ss
            W[0, :, last, :] = loc
        if j > 0:                                    # end a coupling on this site
            W[1, :, last, :] = cp[j] * _sl(HGE3)
            W[2, :, last, :] = -cp[j] * (_sl(HEG3) - _sr(HEG3))
            W[3, :, last, :] = -cpc[j] * (_sr(HGE3) - _sl(HGE3))
            W[4, :, last, :] = cpc[j] * _sr(HEG3)
            for e in range(p.n_exp):
                W[5 + 2 * e, :, last, :] = sk[e] * ket_ss
                W[6 + 2 * e, :, last, :] = sb[e] * bra_ss
            W[last, :, last, :] = I
        if 0 < j < N - 1:                            # pass a coupling through
            W[1, :, 1, :] = ph[j] * I
            W[2, :, 2, :] = ph[j] * I
            W[3, :, 3, :] = np.conj(ph[j]) * I
            W[4, :, 4, :] = np.conj(ph[j]) * I
            for e in range(p.n_exp):
                W[5 + 2 * e, :, 5 + 2 * e, :] = lam[j, e] * I
                W[6 + 2 * e, :, 6 + 2 * e, :] = lam[j, e] * I
        if j == 0:
            W = W[0:1].copy()
        elif j == N - 1:
            W[0, :, last, :] = loc
            W = W[:, :, last:last + 1, :].copy()
        mpo.append(W)
    if N == 1:
        mpo[0] = np.zeros((1, 9, 1, 9), dtype=CDT)
        mpo[0][0, :, 0, :] = rydberg_local_l(p, N, con, delt)
    return mpo


def rydberg_rk4_mpos(p: RydbergParams, z, dt):
    """[L1, L2, L3] of the RK4 half step (rydberg_density_matrix_gpu.jl:199-202)."""
    return [rydberg_lindblad_mpo(p, z, 1.0, dt / 2),
            rydberg_lindblad_mpo(p, z, 0.0, 1.0),
            rydberg_lindblad_mpo(p, z, 1.0, dt / 2)]


def identity_vec_mps(N, d=3, site_vec=None):
    """``makemps([jj->id[:]])`` (rydberg_density_matrix_gpu.jl:134): bond-1 MPS of vec(1)."""
    v = np.eye(d, dtype=CDT).reshape(-1, order="F") if site_vec is None else site_vec
    return [v.reshape(1, d * d, 1).copy() for _ in range(N)]


def local_population_vec_mps(N, j, op, d=3):
    """Identity MPS with site j replaced by vec(op) (``measure_excitations!``, :422-431)."""
    mps = identity_vec_mps(N, d)
    mps[j] = np.asarray(op, dtype=CDT).reshape(-1, order="F").reshape(1, d * d, 1).copy()
    return mps


def sum_local_vec_mps(N, op, d=3):
    """Vectorised sum_j op_j (bond dimension 2), so that one ``scal_prod_no_conj``
    with a vectorised density matrix gives sum_j tr(op_j rho) -- the total
    population that the reference accumulates site by site (``e_pop`` / ``s_pop``,
    rydberg_density_matrix_gpu.jl:174-175)."""
    i = np.eye(d, dtype=CDT).reshape(-1, order="F")
    o = np.asarray(op, dtype=CDT).reshape(-1, order="F")
    dd = d * d
    first = np.zeros((1, dd, 2), dtype=CDT)
    first[0, :, 0], first[0, :, 1] = i, o
    last = np.zeros((2, dd, 1), dtype=CDT)
    last[0, :, 0], last[1, :, 0] = o, i
    bulk = np.zeros((2, dd, 2), dtype=CDT)
    bulk[0, :, 0], bulk[0, :, 1], bulk[1, :, 1] = i, o, i
    if N == 1:
        return [(i * 0 + o).reshape(1, dd, 1)]
    return [first] + [bulk.copy() for _ in range(N - 2)] + [last]


def pair_sum_vec_mps(N, op, d=3):
    """Vectorised sum_{i<j} op_i op_j (bond dimension 3: the bond index counts
    the factors of op placed so far, 0, 1 or 2), so that ``scal_prod_no_conj``
    with a vectorised density matrix gives sum_{i<j} tr(op_i op_j rho); with
    ``sum_local_vec_mps`` this gives the equal-time pair correlation
    g2 = 2 tr(rho) sum_{i<j} <op_i op_j> / (sum_i <op_i>)^2 (N >= 2)."""
    if N < 2:
        raise ValueError("pair sums need N >= 2")
    i = np.eye(d, dtype=CDT).reshape(-1, order="F")
    o = np.asarray(op, dtype=CDT).reshape(-1, order="F")
    dd = d * d
    first = np.zeros((1, dd, 3), dtype=CDT)
    first[0, :, 0], first[0, :, 1] = i, o
    last = np.zeros((3, dd, 1), dtype=CDT)
    last[1, :, 0], last[2, :, 0] = o, i
    bulk = np.zeros((3, dd, 3), dtype=CDT)
    bulk[0, :, 0], bulk[0, :, 1], bulk[1, :, 1], bulk[1, :, 2], bulk[2, :, 2] = i, o, i, o, i
    return [first] + [bulk.copy() for _ in range(N - 2)] + [last]
