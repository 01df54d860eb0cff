"""GPU parity of the quantum-jump trajectory driver against the oracle loop (same draws)."""
import numpy as np
import pytest

from mpsamd import models
from mpsamd.trajectory import SpinTrajectory
from oracle import trajectory as OT

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,D,seed,gam_eg,tol", [(6, 8, 3, 0.5, 0.0), (8, 16, 11, 0.0, 0.0),
                                                 (10, 8, 5, 0.3, 1e-13)])
def test_trajectory_matches_oracle(N, D, seed, gam_eg, tol):
    # full bond dimension for N = 6, 8 (every compression exact); N = 10 at D = 8
    # truncates a zero-padded product state, whose rank-deficient QR panels
    # both sides complete the same way only with the same rank cutoff
    dt, steps = 0.05, 30
    print("case", N, D, seed, gam_eg, tol, flush=True)
    import mpsamd as M
    from oracle import mps_oracle as O
    M.default_context().set_qr_rank_tol(tol)
    O.set_qr_rank_tol(tol)
    tr = SpinTrajectory(N, D, seed, dt=dt, gam_eg=gam_eg, init=(0.0, 1.0))
    H = models.spin_rk4_mpos(tr.z, 0.0, dt)
    jr, jl = models.spin_jump_mpos(tr.z)
    ref = OT.OracleSpinTrajectory(N, D, seed, H, jr, jl, dt, gam_eg, init=(0.0, 1.0))
    for i in range(steps):
        p = tr.step()
        pr = ref.step()
        assert abs(p - pr) < 1e-9 * max(1.0, abs(pr)), (i, p, pr)
        assert tr.jumps == ref.jumps, i
    assert len(tr.jumps) >= 1
    assert np.abs(tr.sigma_z() - ref.sigma_z()).max() < 1e-9
    M.default_context().set_qr_rank_tol(0.0)


def test_trajectory_two_streams_matches_oracle():
    # N = 16 at D = 128: the middle sites (128 x 2 x 128) run the RK4 step on
    # three streams; the state is normalised and may jump between steps, so
    # every step starts from environments rebuilt for the new state
    N, D, seed, gam_eg, tol, dt, steps = 16, 128, 7, 0.3, 1e-13, 0.05, 8
    import mpsamd as M
    from oracle import mps_oracle as O
    M.default_context().set_qr_rank_tol(tol)
    O.set_qr_rank_tol(tol)
    tr = SpinTrajectory(N, D, seed, dt=dt, gam_eg=gam_eg, init=(0.0, 1.0))
    H = models.spin_rk4_mpos(tr.z, 0.0, dt)
    jr, jl = models.spin_jump_mpos(tr.z)
    ref = OT.OracleSpinTrajectory(N, D, seed, H, jr, jl, dt, gam_eg, init=(0.0, 1.0))
    try:
        for i in range(steps):
            p = tr.step()
            pr = ref.step()
            print("step", i, p, pr, flush=True)
            assert abs(p - pr) < 1e-9 * max(1.0, abs(pr)), (i, p, pr)
            assert tr.jumps == ref.jumps, i
        assert np.abs(tr.sigma_z() - ref.sigma_z()).max() < 1e-9
    finally:
        M.default_context().set_qr_rank_tol(0.0)


def test_config_spin_chain_n20_d32_100_steps():
    # the CPU-runnable spin-chain case at its own size: N = 20, D = 32, L = 20,
    # Omega = 0.5, Gamma = 1, k = 2 pi, dt = 0.01, 100 steps from |g...g>;
    # the truncated product-state start needs the same rank cutoff on both sides
    N, D, seed, dt, steps, tol = 20, 32, 1, 0.01, 100, 1e-13
    import mpsamd as M
    from oracle import mps_oracle as O
    M.default_context().set_qr_rank_tol(tol)
    O.set_qr_rank_tol(tol)
    try:
        tr = SpinTrajectory(N, D, seed, dt=dt, L=20.0)
        H = models.spin_rk4_mpos(tr.z, 0.0, dt)
        jr, jl = models.spin_jump_mpos(tr.z)
        ref = OT.OracleSpinTrajectory(N, D, seed, H, jr, jl, dt)
        worst = 0.0
        for i in range(steps):
            p = tr.step()
            pr = ref.step()
            worst = max(worst, abs(p - pr))
            assert abs(p - pr) < 1e-9, (i, p, pr)
            assert tr.jumps == ref.jumps, i
        print("max |p_coh - oracle| over 100 steps", worst, "jumps", tr.jumps, flush=True)
        assert np.abs(tr.sigma_z() - ref.sigma_z()).max() < 1e-9
    finally:
        M.default_context().set_qr_rank_tol(0.0)
