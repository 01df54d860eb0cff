"""Error behaviour of the C-ABI on the device: invalid arguments are rejected on
the host with MPS_E_ARG (DimensionMismatch, as the reference's Julia methods
raise it) before any work is queued, the states they name are left as they
were, and the context stays usable: a valid RK4 step afterwards matches the
oracle.

Every invalid call here is caught by a host-side check in capi.hip that runs
before the first launch of its entry point.
"""
import ctypes as C

import numpy as np
import pytest

import mpsamd as M
from mpsamd import _lib, models
from oracle import mps_oracle as O
from oracle import dense as Dn
from helpers import rel, copy_sites

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _last_error():
    return _lib.load().mps_last_error().decode()


def _setup(N=6, D=8, dt=0.01, seed=1):
    z = models.positions(N, 20.0, seed)
    H = models.spin_rk4_mpos(z, 0.0, dt)
    psi, dims = O.mpsrandom(N, D, 2, np.random.default_rng(seed + 100))
    return H, psi, dims


def test_rk4_rejects_small_environments_and_mixed_bonds():
    N, D, dt = 6, 8, 0.01
    H, psi, dims = _setup(N, D, dt)
    wop = [1] + [4] * (N - 1) + [1]
    small = [min(x, 4) for x in dims]          # [1, 2, 4, 4, 4, 2, 1]
    assert small != dims
    gH = [M.MPO(h) for h in H]
    g = M.MPS.from_host(psi)
    g1, g2, g3 = (M.MPS.from_host(psi) for _ in range(3))
    e1, e2, e3 = (M.build_env(dims, dims) for _ in range(3))
    eop = M.build_env(dims, dims, wop)
    before = g.to_host()

    print("case: op environment built for smaller bonds", flush=True)
    with pytest.raises(M.DimensionMismatch) as ex:
        M.RK4stp_apply_H_half_(dt, gH, g, g1, g2, g3, e1, e2, e3, M.build_env(small, small, wop))
    assert "environment slot 3" in str(ex.value)

    print("case: overlap environment built for smaller bonds", flush=True)
    with pytest.raises(M.DimensionMismatch):
        M.RK4stp_apply_H_half_(dt, gH, g, g1, g2, g3, e1, M.build_env(small, small), e3, eop)

    print("case: a stage state with other bonds", flush=True)
    odd = M.MPS(N, 2, small)
    with pytest.raises(M.DimensionMismatch) as ex:
        M.RK4stp_apply_H_half_(dt, gH, g, odd, g2, g3, e1, e2, e3, eop)
    assert "bond dimensions" in str(ex.value)

    print("case: overlap environment in place of the operator one", flush=True)
    with pytest.raises(M.DimensionMismatch) as ex:
        M.RK4stp_apply_H_half_(dt, gH, g, g1, g2, g3, e1, e2, e3, M.build_env(dims, dims))
    assert "operator environment expected" in _last_error()

    print("case: MPO of another length", flush=True)
    short = [M.MPO(h) for h in models.spin_rk4_mpos(models.positions(N - 1, 20.0, 1), 0.0, dt)]
    with pytest.raises(M.DimensionMismatch):
        M.RK4stp_apply_H_half_(dt, short, g, g1, g2, g3, e1, e2, e3, eop)

    # nothing was written
    after = g.to_host()
    for a, b in zip(after, before):
        assert np.array_equal(a, b)

    # the context is still usable: one valid step equals the oracle's
    print("case: valid step after the rejected calls", flush=True)
    M.RK4stp_apply_H_half_(dt, gH, g, g1, g2, g3, e1, e2, e3, eop)
    n_dev = M.normalize_lo_(g)
    ref = copy_sites(psi)
    oe = [O.build_env(dims, dims) for _ in range(3)]
    O.RK4stp_apply_H_half_(dt, H, ref, copy_sites(psi), copy_sites(psi), copy_sites(psi), *oe,
                           O.build_env_op(dims, dims, wop))
    n_ref = O.normalize_lo_(ref)
    assert abs(n_dev - n_ref) < TOL * abs(n_ref)
    assert rel(Dn.mps_to_dense(g.to_host()), Dn.mps_to_dense(ref)) < TOL


def test_sweeps_reject_mismatched_bonds_and_environments():
    N, D = 6, 8
    rng = np.random.default_rng(7)
    H, psi, dims = _setup(N, D)
    small = [min(x, 4) for x in dims]
    wop = [1] + [4] * (N - 1) + [1]
    gH = M.MPO(H[0])
    g = M.MPS.from_host(psi)
    odd = M.MPS(N, 2, small)

    print("case: compress_var_apply_H_left with a state of other bonds", flush=True)
    with pytest.raises(M.DimensionMismatch):
        M.compress_var_apply_H_left_(M.MPS.from_host(psi), M.build_env(dims, dims, wop), odd, g, gH)

    print("case: compress_var_apply_H with a small operator environment", flush=True)
    with pytest.raises(M.DimensionMismatch):
        M.compress_var_apply_H_(M.MPS.from_host(psi), M.build_env(small, small, wop), g, g, gH, 1)

    print("case: compress_var with an environment smaller than the pair's bonds", flush=True)
    with pytest.raises(M.DimensionMismatch):
        M.compress_var_(M.MPS.from_host(psi), g, M.build_env(small, small), 1)

    print("case: compress_sum_var over states of other bonds", flush=True)
    other, _ = M.mpsrandom(N, 4, 2, rng)
    with pytest.raises(M.DimensionMismatch):
        M.compress_sum_var(g, [0.5, 0.5], [g, other], 1)

    print("case: compress_sum_var_apply_H_right with one overlap term of other bonds", flush=True)
    with pytest.raises(M.DimensionMismatch):
        M.compress_sum_var_apply_H_right_(M.MPS.from_host(psi), [M.build_env(dims, dims)],
                                          M.build_env(dims, dims, wop), g, [1.0, 0.5], [other, g], gH)

    print("case: Arnoldi without Krylov states", flush=True)
    with pytest.raises(M.DimensionMismatch):
        M.Arnoldi_expm_(gH, [], 0, None, M.build_env(dims, dims, wop))

    # g untouched by the rejected calls
    for a, b in zip(g.to_host(), psi):
        assert rel(a, b) < 1e-15


def test_container_and_setting_errors():
    N, D = 5, 4
    H, psi, dims = _setup(N, D)
    g = M.MPS.from_host(psi)
    ctx = M.default_context()
    lib = _lib.load()

    print("case: site index and shape checks", flush=True)
    buf = np.zeros(2 * 64)
    pd = buf.ctypes.data_as(C.POINTER(C.c_double))
    assert lib.mps_state_upload_site(g.h, N, pd, 1, 2, 1) == _lib.MPS_E_ARG
    assert "site out of range" in _last_error()
    assert lib.mps_state_upload_site(g.h, 1, pd, 1, 2, 1) == _lib.MPS_E_ARG
    assert "shape mismatch" in _last_error()
    assert lib.mps_state_download_site(g.h, -1, pd, 1, 2, 1) == _lib.MPS_E_ARG
    with pytest.raises(M.DimensionMismatch):
        M.apply_site_operator_(g, N, np.eye(2, dtype=complex))

    print("case: container creation", flush=True)
    out = C.c_void_p()
    d = (C.c_int64 * 1)(2)
    dm = (C.c_int64 * 2)(1, 1)
    assert lib.mps_state_create(ctx.h, 0, d, dm, C.byref(out)) == _lib.MPS_E_ARG
    bad = (C.c_int64 * 3)(1, 0, 1)
    d2 = (C.c_int64 * 2)(2, 2)
    assert lib.mps_state_create(ctx.h, 2, d2, bad, C.byref(out)) == _lib.MPS_E_ARG
    assert "dims must be >= 1" in _last_error()

    print("case: settings", flush=True)
    assert lib.mps_ctx_set_precision(ctx.h, 7) == _lib.MPS_E_ARG
    assert lib.mps_ctx_set_qr_rank_tol(ctx.h, 1.5) == _lib.MPS_E_ARG
    assert lib.mps_ctx_set_qr_rank_tol(ctx.h, -1e-3) == _lib.MPS_E_ARG
    with pytest.raises(M.DimensionMismatch):
        M.compressionSVD_(M.MPS.from_host(psi), 0)

    print("case: state copy between chains of other lengths", flush=True)
    shorter, _ = M.mpsrandom(N - 1, D, 2, np.random.default_rng(3))
    with pytest.raises(M.DimensionMismatch):
        g.copy_from(shorter)

    # the state and the context settings are as before
    for a, b in zip(g.to_host(), psi):
        assert rel(a, b) < 1e-15
    assert abs(M.norm_lo(g) - O.norm_lo(psi)) < 1e-12
