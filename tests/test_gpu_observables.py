"""Photon-field observables and the photon-detection projection on the device
against the oracle (rydberg_density_matrix_gpu.jl:135-151, 237-245;
rydberg_density_matrix_gpu_g2.jl:117-131, 213): I_r, I2_r, I_l, I2_l through
scal_prod_no_conj with the bond-4/16 operator MPS, and
compress_var_apply_H!(rho, envproj, rho, rho1, ERProj, 4) (output and initial
guess the same state, as the reference calls it), at 1e-10."""
import numpy as np
import pytest

import mpsamd as M
from mpsamd import models as MD
from mpsamd import observables as OB
from oracle import mps_oracle as O
from helpers import copy_sites, rel

pytestmark = pytest.mark.gpu


def _evolved(N, D, steps=3, seed=4):
    # a density matrix a few RK4 steps from the ground state (both sides: oracle)
    p = MD.RydbergParams(n_exp=2)
    z = MD.positions(N, float(N), seed)
    L = MD.rydberg_rk4_mpos(p, z, 0.05)
    O.set_qr_rank_tol(1e-13)
    rho, dims = O.mpsgroundstate(N, D, 9)
    r = [copy_sites(rho) for _ in range(3)]
    e = [O.build_env(dims, dims) for _ in range(3)]
    eop = O.build_env_op(dims, dims, [1] + [p.w] * (N - 1) + [1])
    for _ in range(steps):
        O.RK4stp_apply_H_half_(0.05, L, rho, *r, *e, eop)
    O.set_qr_rank_tol(0.0)
    return z, rho, dims


def test_photon_intensities_and_g2():
    N, D = 6, 16
    z, rho, dims = _evolved(N, D)
    obs = OB.photon_observables(z, f_in=0.3)
    g = M.MPS.from_host(rho)
    scale = abs(O.scal_prod_no_conj(obs["IR"], rho))   # the largest of the four
    assert scale > 1e-3
    for key in ("IR", "I2R", "IL", "I2L"):
        want = O.scal_prod_no_conj(obs[key], rho)
        got = M.scal_prod_no_conj(M.MPS.from_host(obs[key]), g)
        print(key, got, want, flush=True)
        assert abs(want) > 1e-12
        # relative to the intensities' scale: I2_l of this weakly excited
        # state is a sum of terms of that scale cancelling to ~1e-9
        assert abs(got - want) < 1e-10 * max(abs(want), scale)
    Ir = M.scal_prod_no_conj(M.MPS.from_host(obs["IR"]), g)
    I2r = M.scal_prod_no_conj(M.MPS.from_host(obs["I2R"]), g)
    want = O.scal_prod_no_conj(obs["I2R"], rho) / O.scal_prod_no_conj(obs["IR"], rho) ** 2
    assert abs(OB.g2(Ir, I2r) - want) < 1e-10 * abs(want)


@pytest.mark.parametrize("f_amp", [0.0, 0.25])
def test_er_projection_four_sweeps(f_amp, rng):
    N, D = 6, 16
    z, _, dims = _evolved(N, D)
    rho, dims = O.mpsrandom(N, D, 9, rng)
    P = OB.er_projection_mpo(z, f_amp=f_amp)
    wd = [1] + [4] * (N - 1) + [1]
    ref = copy_sites(rho)
    O.compress_var_apply_H_(ref, O.build_env_op(dims, dims, wd), ref, copy_sites(rho), P, 4)
    g = M.MPS.from_host(rho)
    g1 = M.MPS.from_host(rho)
    M.compress_var_apply_H_(g, M.build_env(dims, dims, wd), g, g1, M.MPO(P), 4)
    got = g.to_host()
    for j in range(N):
        assert rel(got[j], ref[j]) < 1e-10, j
