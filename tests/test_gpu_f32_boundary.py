"""ComplexF32 host buffers at the boundary (the reference's GPU examples run
TN = Complex{Float32}, rydberg_density_matrix_gpu.jl:124-125): uploads widen
exactly, downloads round to nearest, and an RK4 step of complex64 inputs
(state and MPOs) equals the FP64 oracle on the same rounded inputs to 1e-10,
and the FP64-input result to FP32 tolerance."""
import numpy as np
import pytest

import mpsamd as M
from mpsamd import models
from oracle import mps_oracle as O
from helpers import copy_sites, rel

pytestmark = pytest.mark.gpu


def test_f32_site_transfers_round_trip(rng):
    a, dims = O.mpsrandom(5, 12, 9, rng)
    a32 = [s.astype(np.complex64) for s in a]
    g = M.MPS.from_host(a32)                       # ComplexF32 upload
    for j in range(5):
        assert np.array_equal(g.site(j), a32[j].astype(np.complex128))
    h = M.MPS.from_host(a)                         # FP64 upload, ComplexF32 download
    for j in range(5):
        assert np.array_equal(h.site_f32(j), a[j].astype(np.complex64))


def test_rk4_with_complex64_inputs():
    N, D, dt = 6, 16, 0.01
    p = models.RydbergParams(n_exp=2)
    z = models.positions(N, float(N), 3)
    L = models.rydberg_rk4_mpos(p, z, dt)
    rho, dims = O.mpsrandom(N, D, 9, np.random.default_rng(8))
    wop = [1] + [p.w] * (N - 1) + [1]
    L32 = [[s.astype(np.complex64) for s in h] for h in L]
    rho32 = [s.astype(np.complex64) for s in rho]

    def oracle(Lh, r0):
        r0 = [s.astype(np.complex128) for s in r0]
        Lh = [[s.astype(np.complex128) for s in h] for h in Lh]
        r = [copy_sites(r0) for _ in range(3)]
        O.RK4stp_apply_H_half_(dt, Lh, r0, *r, *[O.build_env(dims, dims) for _ in range(3)],
                               O.build_env_op(dims, dims, wop))
        return r0

    ref32 = oracle(L32, rho32)
    ref64 = oracle(L, rho)
    g = M.MPS.from_host(rho32)
    gr = [M.MPS.from_host(rho32) for _ in range(3)]
    gL = [M.MPO(h) for h in L32]                  # complex64 sites: ComplexF32 upload
    M.RK4stp_apply_H_half_(dt, gL, g, *gr, *[M.build_env(dims, dims) for _ in range(3)],
                           M.build_env(dims, dims, wop))
    got = g.to_host()
    for j in range(N):
        assert rel(got[j], ref32[j]) < 1e-10, j
        assert rel(got[j], ref64[j]) < 1e-5, j
