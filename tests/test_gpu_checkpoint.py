"""Checkpoint / resume on the device: a run interrupted at step k, written to
a snapshot file, reloaded into freshly built objects on a new context and
continued, equals the uninterrupted run bit for bit (the reference's temp-file
restart, rydberg_density_matrix_gpu.jl:251-255, 273-316 and
rydberg_pollution_density_matrix_gpu_restart.jl:151-189).

* a density matrix at D = 64 (three-stream RK4 path), FP64 and the FP32
  form with refinement;
* a quantum-jump trajectory with jumps (state, draws, records);
* in-flight served requests (two spin trajectories and a density matrix),
  snapshotted between ticks, their solves released, restored and finished."""
import numpy as np
import pytest

import mpsamd as M
from mpsamd import checkpoint as CK
from mpsamd import serving as S
from mpsamd.trajectory import SpinTrajectory

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,nref", [("fp64", 0), ("fp32", 2)])
def test_density_matrix_resume_bit_exact(prec, nref, tmp_path):
    N, D, k, m = 6, 64, 3, 3
    print("case density matrix resume", prec, nref, flush=True)

    def evolution():
        ctx = M.Context(0)
        ctx.set_qr_rank_tol(S.QR_RANK_TOL)
        ctx.set_precision(prec)
        ctx.set_refine_sweeps(nref)
        return S.RydbergEvolution(N, D, 5, n_exp=4, ctx=ctx)

    a = evolution()
    for _ in range(k + m):
        a.launch_coherent()
    want = a.rho.to_host()
    b = evolution()
    for _ in range(k):
        b.launch_coherent()
    path = str(tmp_path / "rho.npz")
    b.checkpoint(path, step=k)
    c = evolution()
    meta = c.resume(path)
    assert meta["step"] == k
    for _ in range(m):
        c.launch_coherent()
    got = c.rho.to_host()
    for j in range(N):
        assert np.array_equal(got[j], want[j]), j


def test_spin_trajectory_resume_bit_exact(tmp_path):
    N, D, seed, dt, k, m = 10, 8, 5, 0.05, 12, 18
    ctx = M.default_context()
    ctx.set_qr_rank_tol(1e-13)
    try:
        a = SpinTrajectory(N, D, seed, dt=dt, gam_eg=0.3, init=(0.0, 1.0))
        for _ in range(k + m):
            a.step()
        b = SpinTrajectory(N, D, seed, dt=dt, gam_eg=0.3, init=(0.0, 1.0))
        for _ in range(k):
            b.step()
        path = str(tmp_path / "traj.npz")
        b.checkpoint(path)
        c = SpinTrajectory(N, D, seed, dt=dt, gam_eg=0.3, init=(0.0, 1.0))
        c.resume(path)
        for _ in range(m):
            c.step()
    finally:
        ctx.set_qr_rank_tol(0.0)
    assert len(a.jumps) >= 1 and a.jumps == c.jumps and a.p_coh == c.p_coh
    for x, y in zip(a.A.to_host(), c.A.to_host()):
        assert np.array_equal(x, y)


def test_served_requests_snapshot_and_restore(tmp_path):
    spec = [(0, "spin", 12, 16, 10, 3), (1, "spin", 10, 8, 7, 4), (2, "rydberg", 6, 16, 8, 5)]
    full = [S.Request(*s) for s in spec]
    S.run_batch(full)
    part = [S.Request(*s) for s in spec]
    S.run_batch(part, max_ticks=4)
    inflight = [r for r in part if r.solve is not None]
    assert len(inflight) == 3
    S.snapshot_requests(inflight, str(tmp_path / "snap"))
    for r in inflight:
        S.release(r.solve)
        r.solve = None
    back = S.restore_requests(str(tmp_path / "snap"))
    assert [r.left for r in back] == [6, 3, 4]
    S.run_batch(back, resumed=True)
    for a, b in zip(full, sorted(back, key=lambda r: r.rid)):
        assert np.array_equal(a.result, b.result), a.rid
