"""Lockstep ensemble of trajectories (one batched context) against the same
trajectories run one by one (plain contexts), and against the oracle loop."""
import numpy as np
import pytest

import mpsamd as M
from mpsamd import models
from mpsamd.trajectory import SpinEnsemble, SpinTrajectory
from oracle import trajectory as OT

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,D,pos_seed,seeds,gam_eg,tol", [(8, 16, 4, [11, 12, 13], 0.5, 0.0),
                                                          (10, 8, 9, [1, 2, 3, 4, 5], 0.0, 1e-13)])
def test_ensemble_matches_single_trajectories(N, D, pos_seed, seeds, gam_eg, tol):
    # decaying from |e...e>: jumps of every kind within the first steps; N = 10
    # at D = 8 truncates, with the rank cutoff on both sides
    dt, steps = 0.05, 20
    print("case", N, D, pos_seed, seeds, gam_eg, tol, flush=True)
    ens = SpinEnsemble(N, D, pos_seed, seeds, dt=dt, gam_eg=gam_eg, init=(0.0, 1.0), qr_rank_tol=tol)
    singles = []
    for s in seeds:
        ctx = M.Context(0)
        ctx.set_qr_rank_tol(tol)
        singles.append((ctx, SpinTrajectory(N, D, s, dt=dt, gam_eg=gam_eg, init=(0.0, 1.0), ctx=ctx,
                                            pos_seed=pos_seed)))
    for i in range(steps):
        p = ens.step()
        for b, (_, tr) in enumerate(singles):
            pb = tr.step()
            assert abs(p[b] - pb) < 1e-11, (i, b, p[b], pb)
            assert ens.jumps[b] == tr.jumps, (i, b)
    assert sum(len(j) for j in ens.jumps) >= len(seeds)
    for b, (_, tr) in enumerate(singles):
        assert np.abs(ens.sigma_z(b) - tr.sigma_z()).max() < 1e-11
    ens.close()
    for ctx, tr in singles:
        for v in list(vars(tr).values()):
            for o in (v if isinstance(v, (list, tuple)) else [v]):
                if isinstance(o, (M.MPS, M.MPO, M.Env)):
                    o.__del__()
        ctx.close()


def test_ensemble_member_matches_oracle():
    # full bond dimension (N = 6, D = 8): every compression exact
    N, D, pos_seed, seeds, dt, steps = 6, 8, 3, [21, 22], 0.05, 15
    ens = SpinEnsemble(N, D, pos_seed, seeds, dt=dt, gam_eg=0.5, init=(0.0, 1.0))
    H = models.spin_rk4_mpos(ens.z, 0.0, dt)
    jr, jl = models.spin_jump_mpos(ens.z)
    refs = [OT.OracleSpinTrajectory(N, D, s, H, jr, jl, dt, 0.5, init=(0.0, 1.0)) for s in seeds]
    for i in range(steps):
        p = ens.step()
        for b, ref in enumerate(refs):
            pr = ref.step()
            assert abs(p[b] - pr) < 1e-9, (i, b, p[b], pr)
            assert ens.jumps[b] == ref.jumps, (i, b)
    for b, ref in enumerate(refs):
        assert np.abs(ens.sigma_z(b) - ref.sigma_z()).max() < 1e-9
    ens.close()


def test_batched_context_refuses_single_copy_calls():
    ens = SpinEnsemble(6, 8, 1, [1, 2], dt=0.05)
    with pytest.raises(M.MPSError) as e:
        M.scal_prod(ens.A, ens.A)
    assert e.value.code == -4          # MPS_E_STATE
    # the same call on a view of one copy works
    v = ens.Av[0]
    assert abs(M.scal_prod(v, v) - 1.0) < 1e-12
    ens.close()


def test_normalize_lo_batch_per_copy():
    # normalize_lo! of every lockstep copy in one call (orthonormalisation.jl:97-121):
    # copies hold different left-canonical states scaled differently; each
    # returned norm^2 and each normalised copy equal the oracle's on that copy
    from oracle import mps_oracle as O
    from helpers import copy_sites
    from mpsamd.trajectory import _state_bytes
    N, D, d, nb = 6, 8, 3, 3
    rng = np.random.default_rng(21)
    states = []
    for b in range(nb):
        s, dims = O.mpsrandom(N, D, d, rng)
        s[-1] = s[-1] * (0.5 + b)
        states.append(s)
    ctx = M.Context.batched(nb, _state_bytes(dims, d) + 4096, M.batch_workspace_bytes(nb, D, d, 1), 0)
    comp = ctx.companion()
    g = M.MPS(N, d, dims, ctx)
    views = [g.view(b, comp) for b in range(nb)]
    for v, s in zip(views, states):
        v.upload(s)
    norms = M.normalize_lo_batch(g)
    for b in range(nb):
        ref = copy_sites(states[b])
        n_ref = O.normalize_lo_(ref)
        assert abs(norms[b] - n_ref) < 1e-12 * n_ref, b
        got = views[b].to_host()
        assert max(np.abs(x - y).max() for x, y in zip(got, ref)) < 1e-12, b
    for v in views:
        v.__del__()
    g.__del__()
    comp.close()
    ctx.close()
