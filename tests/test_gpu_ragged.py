"""Ragged pools: quantum-jump trajectories of chains of different lengths in
one lockstep pool, advanced by one RK4 call whose every site-slot launch
covers all chains present there.  Checked against the same trajectories in
uniform pools of the same size (bit for bit: same kernels, same GEMM plans),
alone on plain contexts, and the oracle loop."""
import numpy as np
import pytest

import mpsamd as M
from mpsamd import models
from mpsamd.trajectory import RaggedSpinEnsemble, SpinEnsemble, SpinTrajectory
from oracle import mps_oracle as O
from oracle import trajectory as OT

pytestmark = pytest.mark.gpu

# decaying from |e...e> with local emission: jumps of every kind early on;
# D = 4 truncates every chain, so the rank cutoff is set on every side
NS = [12, 9, 16, 10, 13]
POS = [3, 4, 5, 6, 7]
SEEDS = [31, 32, 33, 34, 35]
KW = dict(dt=0.05, gam_eg=0.5, init=(0.0, 1.0), qr_rank_tol=1e-13)


def _close_single(ctx, tr):
    for v in list(vars(tr).values()):
        for o in (v if isinstance(v, (list, tuple)) else [v]):
            if isinstance(o, (M.MPS, M.MPO, M.Env)):
                o.__del__()
    ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ragged_pool_equals_uniform_pools_bit_for_bit(precision):
    D, steps = 4, 12
    print("case ragged vs uniform", precision, NS, flush=True)
    rag = RaggedSpinEnsemble(NS, D, POS, SEEDS, precision=precision, **KW)
    assert rag.order == [2, 4, 0, 3, 1]
    unis = []
    for i, n in enumerate(NS):
        # trajectory i as copy 0 of a uniform pool of as many copies
        seeds = [SEEDS[i]] + [100 + k for k in range(len(NS) - 1)]
        u = SpinEnsemble(n, D, POS[i], seeds, **KW)
        u.ctx.set_precision(precision)
        unis.append(u)
    for s in range(steps):
        p = rag.step()
        for i, u in enumerate(unis):
            pu = u.step()
            assert p[i] == pu[0], (s, i, p[i], pu[0])
            assert rag.jumps[i] == u.jumps[0], (s, i)
    assert sum(len(j) for j in rag.jumps) >= len(NS)
    for i, u in enumerate(unis):
        a, b = rag.state(i), u.Av[0].to_host()
        assert len(a) == NS[i]
        assert max(np.abs(x - y).max() for x, y in zip(a, b)) == 0.0, i
        assert np.array_equal(rag.sigma_z(i), u.sigma_z(0))
        u.close()
    rag.close()


@pytest.mark.parametrize("D", [64, 128])
def test_ragged_pool_equals_uniform_pools_at_benchmark_lengths(D):
    # chains of 40-120 sites (the uneven batch's range) at the batch's bond
    # dimensions, FP32 contractions and the batch leg's settings, in one
    # ragged pool against the same trajectories in uniform pools of as many
    # copies: bit for bit, as at the small sizes above
    Ns, pos, seeds, steps = [120, 57, 83, 40], [11, 12, 13, 14], [51, 52, 53, 54], 2
    kw = dict(dt=0.01, qr_rank_tol=1e-13)
    print("case ragged vs uniform at benchmark lengths", D, Ns, flush=True)
    rag = RaggedSpinEnsemble(Ns, D, pos, seeds, precision="fp32", **kw)
    unis = []
    for i, n in enumerate(Ns):
        u = SpinEnsemble(n, D, pos[i], [seeds[i]] + [200 + k for k in range(len(Ns) - 1)], **kw)
        u.ctx.set_precision("fp32")
        unis.append(u)
    for s in range(steps):
        p = rag.step()
        for i, u in enumerate(unis):
            pu = u.step()
            assert p[i] == pu[0], (s, i, p[i], pu[0])
            assert rag.jumps[i] == u.jumps[0], (s, i)
    for i, u in enumerate(unis):
        assert np.array_equal(rag.sigma_z(i), u.sigma_z(0)), i
        u.close()
    rag.close()


def test_ragged_pool_matches_single_trajectories():
    D, steps = 4, 12
    print("case ragged vs single", NS, flush=True)
    rag = RaggedSpinEnsemble(NS, D, POS, SEEDS, **KW)
    singles = []
    for i, n in enumerate(NS):
        ctx = M.Context(0)
        ctx.set_qr_rank_tol(KW["qr_rank_tol"])
        singles.append((ctx, SpinTrajectory(n, D, SEEDS[i], dt=KW["dt"], gam_eg=KW["gam_eg"], init=KW["init"],
                                            ctx=ctx, pos_seed=POS[i])))
    for s in range(steps):
        p = rag.step()
        for i, (_, tr) in enumerate(singles):
            pt = tr.step()
            assert abs(p[i] - pt) < 1e-11, (s, i, p[i], pt)
            assert rag.jumps[i] == tr.jumps, (s, i)
    for i, (_, tr) in enumerate(singles):
        assert np.abs(rag.sigma_z(i) - tr.sigma_z()).max() < 1e-11
    for ctx, tr in singles:
        _close_single(ctx, tr)
    rag.close()


def test_ragged_pool_matches_oracle():
    # full bond dimension (D = 8 holds chains of 6 and 7 sites exactly): every
    # compression is exact, so the oracle and the device agree without a cutoff
    Ns, pos, seeds, D, steps = [7, 6, 7, 6], [1, 2, 3, 4], [41, 42, 43, 44], 8, 15
    kw = dict(KW, qr_rank_tol=0.0)
    print("case ragged vs oracle", Ns, flush=True)
    try:
        rag = RaggedSpinEnsemble(Ns, D, pos, seeds, **kw)
        refs = []
        for i, n in enumerate(Ns):
            z = models.positions(n, float(n), pos[i])
            refs.append(OT.OracleSpinTrajectory(n, D, seeds[i], models.spin_rk4_mpos(z, 0.0, KW["dt"]),
                                                *models.spin_jump_mpos(z), KW["dt"], KW["gam_eg"],
                                                init=KW["init"]))
        for s in range(steps):
            p = rag.step()
            for i, ref in enumerate(refs):
                pr = ref.step()
                assert abs(p[i] - pr) < 1e-9, (s, i, p[i], pr)
                assert rag.jumps[i] == ref.jumps, (s, i)
        assert sum(len(j) for j in rag.jumps) >= len(Ns)
        for i, ref in enumerate(refs):
            assert np.abs(rag.sigma_z(i) - ref.sigma_z()).max() < 1e-9
        rag.close()
    finally:
        O.set_qr_rank_tol(0.0)


def test_ragged_pool_errors():
    # chains must be sorted longest first, and only the per-slot sweeps take the pool
    from mpsamd.core import _dims_rule
    dims = _dims_rule(12, 4, 2)
    from mpsamd.trajectory import slab_sizes
    Hs = models.spin_rk4_mpos(models.positions(12, 12.0, 0), 0.0, 0.05)
    cb, ws, _ = slab_sizes(dims, 2, Hs, 5, 3)
    ctx = M.Context.batched(3, cb, ws)
    with pytest.raises(M.MPSError) as e:
        ctx.set_ragged([12, 8, 10])
    assert e.value.code == -1
    ctx.set_ragged([12, 10, 8])
    A = M.MPS(12, 2, dims, ctx)
    with pytest.raises(M.MPSError) as e:
        M.leftorthonormalizationQR_(A)
    assert e.value.code == -4
    with pytest.raises(M.MPSError) as e:
        M.mpsproductstate(12, 4, 2, np.array([1.0, 0.0]), ctx)
    assert e.value.code == -4
    A.__del__()
    ctx.close()
    # a chain whose two parts would meet at an edge bond is refused
    with pytest.raises(ValueError):
        RaggedSpinEnsemble([20, 6], 16, [1, 2], [1, 2])
