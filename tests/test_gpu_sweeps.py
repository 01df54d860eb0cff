"""GPU parity of the sweeps, RK4, gauge moves and scalar products against the oracle.

Inputs are seeded random left-canonical states (mpsrandom), product states and
the spin / Rydberg model MPOs; the device result must match the FP64 oracle
to 1e-10 relative (observables, dense states) on the same inputs.
"""
import numpy as np
import pytest

import mpsamd as M
from mpsamd import models
from oracle import mps_oracle as O
from oracle import dense as Dn
from helpers import rel, copy_sites, random_mpo, crandn

pytestmark = pytest.mark.gpu
TOL = 1e-10


def dev_state(sites):
    return M.MPS.from_host(sites)


def dev_sites(m):
    return m.to_host()


def env_dims(dims):
    return list(dims)


@pytest.mark.parametrize("N,D,d", [(6, 8, 2), (8, 16, 2), (5, 9, 3)])
def test_compress_var_apply_H_left(N, D, d, rng):
    print("case", N, D, d, flush=True)
    psi, dims = O.mpsrandom(N, D, d, rng)
    H = random_mpo(rng, N, d, 3)
    A0 = copy_sites(psi)
    envop = O.build_env_op(dims, dims, [1] + [3] * (N - 1) + [1])
    O.compress_var_apply_H_left_(A0, envop, psi, psi, H)

    g_psi = dev_state(psi)
    g_A0 = dev_state(psi)
    g_H = M.MPO(H)
    g_env = M.build_env(dims, dims, [1] + [3] * (N - 1) + [1])
    M.compress_var_apply_H_left_(g_A0, g_env, g_psi, g_psi, g_H)
    got = dev_sites(g_A0)
    for j in range(N):
        assert rel(got[j], A0[j]) < TOL, j


@pytest.mark.parametrize("side", ["left", "right"])
@pytest.mark.parametrize("numv", [2, 3, 4])
def test_compress_sum_var_apply_H(side, numv, rng):
    N, D, d, w = 7, 8, 2, 4
    print("case", side, numv, flush=True)
    states = [O.mpsrandom(N, D, d, rng)[0] for _ in range(numv)]
    Ain, dims = O.mpsrandom(N, D, d, rng)
    if side == "right":   # Ain right canonical for the _right sweep
        O.rightorthonormalizationQR_(Ain)
    H = random_mpo(rng, N, d, w)
    coef = list(crandn(rng, numv))
    wd = [1] + [w] * (N - 1) + [1]
    A0 = copy_sites(Ain)
    env = [O.build_env(dims, dims) for _ in range(numv - 1)]
    envop = O.build_env_op(dims, dims, wd)
    fn = O.compress_sum_var_apply_H_left_ if side == "left" else O.compress_sum_var_apply_H_right_
    fn(A0, env, envop, Ain, coef, states, H)

    g_states = [dev_state(s) for s in states]
    g_Ain = dev_state(Ain)
    g_A0 = dev_state(Ain)
    g_env = [M.build_env(dims, dims) for _ in range(numv - 1)]
    g_envop = M.build_env(dims, dims, wd)
    gfn = M.compress_sum_var_apply_H_left_ if side == "left" else M.compress_sum_var_apply_H_right_
    gfn(g_A0, g_env, g_envop, g_Ain, coef, g_states, M.MPO(H))
    got = dev_sites(g_A0)
    for j in range(N):
        assert rel(got[j], A0[j]) < TOL, j


def test_sum_var_requires_an_overlap_term(rng):
    # the reference indexes A[1] unconditionally (compression.jl:918, 1036): numv = 1 is a BoundsError
    N, D, d = 5, 4, 2
    s, dims = O.mpsrandom(N, D, d, rng)
    H = random_mpo(rng, N, d, 2)
    wd = [1] + [2] * (N - 1) + [1]
    with pytest.raises(IndexError):
        O.compress_sum_var_apply_H_left_(copy_sites(s), [], O.build_env_op(dims, dims, wd), s, [1.0], [s], H)
    g = dev_state(s)
    with pytest.raises(M.DimensionMismatch):
        M.compress_sum_var_apply_H_left_(dev_state(s), [], M.build_env(dims, dims, wd), g, [1.0], [g], M.MPO(H))


def _spin_setup(N, D, dt, init, seed=1):
    z = models.positions(N, 20.0, seed)
    H = models.spin_rk4_mpos(z, 0.0, dt)
    if init == "product":
        psi, dims = O.mpsproductstate(N, D, 2, np.array([1, 0], dtype=complex))
    else:
        psi, dims = O.mpsrandom(N, D, 2, np.random.default_rng(seed + 100))
    return z, H, psi, dims


def _oracle_rk4(H, psi, dims, dt, steps, wop, eleven=False):
    psi = copy_sites(psi)
    envop = O.build_env_op(dims, dims, wop)
    e1, e2, e3 = (O.build_env(dims, dims) for _ in range(3))
    p1, p2, p3 = copy_sites(psi), copy_sites(psi), copy_sites(psi)
    norms = []
    for _ in range(steps):
        if eleven:
            out = copy_sites(psi)
            O.RK4stp_apply_H_half_(dt, H, psi, p1, p2, p3, e1, e2, e3, envop, psi0=out)
            psi = out
        else:
            O.RK4stp_apply_H_half_(dt, H, psi, p1, p2, p3, e1, e2, e3, envop)
        norms.append(O.normalize_lo_(psi))
    return psi, norms


def _device_rk4(H, psi, dims, dt, steps, wop, eleven=False, graphs=True):
    ctx = M.default_context()
    ctx.set_graphs(graphs)
    gH = [M.MPO(h) for h in H]
    g = dev_state(psi)
    g1, g2, g3 = dev_state(psi), dev_state(psi), dev_state(psi)
    e1, e2, e3 = (M.build_env(dims, dims) for _ in range(3))
    eop = M.build_env(dims, dims, wop)
    out = dev_state(psi) if eleven else None
    norms = []
    for _ in range(steps):
        if eleven:
            M.RK4stp_apply_H_half_(dt, gH, out, g, g1, g2, g3, e1, e2, e3, eop)
            g.copy_from(out)
        else:
            M.RK4stp_apply_H_half_(dt, gH, g, g1, g2, g3, e1, e2, e3, eop)
        norms.append(M.normalize_lo_(g))
    ctx.set_graphs(True)
    return g, norms


@pytest.mark.parametrize("eleven", [False, True])
@pytest.mark.parametrize("graphs", [False, True])
def test_rk4_random_state(eleven, graphs):
    N, D, dt, steps = 8, 8, 0.01, 4
    print("case", eleven, graphs, flush=True)
    z, H, psi, dims = _spin_setup(N, D, dt, "random")
    wop = [1] + [4] * (N - 1) + [1]
    ref, rn = _oracle_rk4(H, psi, dims, dt, steps, wop, eleven)
    g, gn = _device_rk4(H, psi, dims, dt, steps, wop, eleven, graphs)
    assert np.allclose(gn, rn, rtol=TOL, atol=0)
    got = g.to_host()
    assert rel(Dn.mps_to_dense(got), Dn.mps_to_dense(ref)) < TOL


def test_rk4_product_state_full_bond_dimension():
    # N=8, D=16 = full dimension: the one-site projection is exact and the
    # zero-padded QR completions are immaterial.
    N, D, dt, steps = 8, 16, 0.01, 5
    z, H, psi, dims = _spin_setup(N, D, dt, "product")
    wop = [1] + [4] * (N - 1) + [1]
    ref, rn = _oracle_rk4(H, psi, dims, dt, steps, wop)
    g, gn = _device_rk4(H, psi, dims, dt, steps, wop)
    assert np.allclose(gn, rn, rtol=TOL, atol=0)
    assert rel(Dn.mps_to_dense(g.to_host()), Dn.mps_to_dense(ref)) < TOL


def test_rk4_product_state_truncated_with_rank_cutoff():
    # Truncated bond dimension from a zero-padded product state: with the
    # rank cutoff 1e-13 both sides complete Q identically (tau = 0 on
    # rounding-level columns) and stay within 1e-10 over 30 steps.
    N, D, dt, steps = 12, 8, 0.01, 30
    z, H, psi, dims = _spin_setup(N, D, dt, "product")
    wop = [1] + [4] * (N - 1) + [1]
    O.set_qr_rank_tol(1e-13)
    ctx = M.default_context()
    ctx.set_qr_rank_tol(1e-13)
    try:
        ref, rn = _oracle_rk4(H, psi, dims, dt, steps, wop)
        g, gn = _device_rk4(H, psi, dims, dt, steps, wop)
    finally:
        ctx.set_qr_rank_tol(0.0)
        O.set_qr_rank_tol(0.0)
    assert np.allclose(gn, rn, rtol=TOL, atol=0)
    ez_ref = _sz(ref)
    ez_dev = M.measure_excitations_(None, g, models.HEE2)
    assert np.abs(ez_dev - ez_ref).max() < TOL


@pytest.mark.parametrize("N,D", [(2, 4), (3, 4), (4, 2)])
def test_rk4_short_chains(N, D):
    # the shortest chains: every slot of the environments is a boundary slot
    # or next to one, and the sweeps run one or two sites
    dt, steps = 0.01, 4
    print("case", N, D, flush=True)
    z, H, psi, dims = _spin_setup(N, D, dt, "random", seed=3)
    wop = [1] + [4] * (N - 1) + [1]
    ref, rn = _oracle_rk4(H, psi, dims, dt, steps, wop)
    g, gn = _device_rk4(H, psi, dims, dt, steps, wop)
    assert np.allclose(gn, rn, rtol=TOL, atol=0)
    assert rel(Dn.mps_to_dense(g.to_host()), Dn.mps_to_dense(ref)) < TOL
    assert np.abs(M.measure_excitations_(None, g, models.HEE2) - _sz(ref)).max() < TOL


def test_rk4_mpos_of_another_context_replaced_between_runs():
    # MPOs owned by a second context: such a call runs eagerly (a graph of it
    # could outlive the MPOs), so MPOs destroyed and re-made with other
    # contents, possibly at the same host address, are the ones used
    N, D, steps = 8, 8, 3
    ctx2 = M.Context(0)
    try:
        for dt in (0.01, 0.02, 0.01):
            print("case dt", dt, flush=True)
            z, H, psi, dims = _spin_setup(N, D, dt, "random")
            wop = [1] + [4] * (N - 1) + [1]
            ref, rn = _oracle_rk4(H, psi, dims, dt, steps, wop)
            gH = [M.MPO(h, ctx2) for h in H]
            g = dev_state(psi)
            g1, g2, g3 = dev_state(psi), dev_state(psi), dev_state(psi)
            e1, e2, e3 = (M.build_env(dims, dims) for _ in range(3))
            eop = M.build_env(dims, dims, wop)
            gn = []
            for _ in range(steps):
                M.RK4stp_apply_H_half_(dt, gH, g, g1, g2, g3, e1, e2, e3, eop)
                gn.append(M.normalize_lo_(g))
            assert np.allclose(gn, rn, rtol=TOL, atol=0)
            assert rel(Dn.mps_to_dense(g.to_host()), Dn.mps_to_dense(ref)) < TOL
            for h in gH:
                h.__del__()
    finally:
        ctx2.close()


def _sz(psi):
    N = len(psi)
    env = np.ones((1, 1), dtype=complex)
    ep = np.zeros(N)
    for n in range(N - 1, -1, -1):
        Aenv = O.prod_LR_32(psi[n], env)
        ep[n] = O.local_exp(Aenv, psi[n], models.HEE2).real
        env = O.update_renv(psi[n], Aenv)
    return ep


def test_rk4_rydberg_density_matrix():
    N, D, dt, steps = 5, 12, 0.005, 3
    p = models.RydbergParams(n_exp=2)
    z = models.positions(N, 5.0, 3)
    L = models.rydberg_rk4_mpos(p, z, dt)
    rho, dims = O.mpsrandom(N, D, 9, np.random.default_rng(5))
    wop = [1] + [p.w] * (N - 1) + [1]
    ref, rn = _oracle_rk4(L, rho, dims, dt, steps, wop)
    g, gn = _device_rk4(L, rho, dims, dt, steps, wop)
    assert np.allclose(gn, rn, rtol=TOL, atol=0)
    I = models.identity_vec_mps(N)
    gI = dev_state(I)
    assert abs(M.scal_prod_no_conj(gI, g) - O.scal_prod_no_conj(I, ref)) < TOL


def test_scalar_products_and_observables(rng):
    N, D, d = 7, 6, 3
    a, dims = O.mpsrandom(N, D, d, rng)
    b, _ = O.mpsrandom(N, D, d, rng)
    H = random_mpo(rng, N, d, 3)
    ga, gb, gH = dev_state(a), dev_state(b), M.MPO(H)
    assert abs(M.scal_prod(ga, gb) - O.scal_prod(a, b)) < TOL
    assert abs(M.scal_prod_no_conj(ga, gb) - O.scal_prod_no_conj(a, b)) < TOL
    assert abs(M.scal_op_prod(ga, gH, gb) - O.scal_op_prod(a, H, b)) < TOL
    op = crandn(rng, (d, d))
    op = op + op.conj().T
    assert np.abs(M.measure_excitations_(None, ga, op) - _local(a, op)).max() < TOL
    M.apply_site_operator_(ga, 3, op)
    O.apply_site_operator_(a[3], op)
    assert rel(ga.site(3), a[3]) < TOL


def _local(psi, op):
    N = len(psi)
    env = np.ones((1, 1), dtype=complex)
    ep = np.zeros(N)
    for n in range(N - 1, -1, -1):
        Aenv = O.prod_LR_32(psi[n], env)
        ep[n] = O.local_exp(Aenv, psi[n], op).real
        env = O.update_renv(psi[n], Aenv)
    return ep


def test_mpo_set_block(rng):
    # mps_mpo_set_block rewrites O[j][wl, :, wr, :] in place (the per-step
    # writes of update_h! / update_L_Ryd!); through <A|O|B> its effect equals
    # the same rewrite of the host MPO in the oracle
    N, D, d, w = 6, 6, 3, 3
    a, _ = O.mpsrandom(N, D, d, rng)
    b, _ = O.mpsrandom(N, D, d, rng)
    H = random_mpo(rng, N, d, w)
    gH = M.MPO(H)
    ga, gb = dev_state(a), dev_state(b)
    before = O.scal_op_prod(a, H, b)
    assert abs(M.scal_op_prod(ga, gH, gb) - before) < TOL * max(1.0, abs(before))
    for j, wl, wr in [(0, 0, 2), (2, 1, 0), (2, 2, 2), (5, 1, 0)]:
        op = crandn(rng, (d, d))
        H[j][wl, :, wr, :] = op
        gH.set_block(j, wl, wr, op)
    after = O.scal_op_prod(a, H, b)
    assert abs(after - before) > 1e-3 * abs(before)
    assert abs(M.scal_op_prod(ga, gH, gb) - after) < TOL * max(1.0, abs(after))


@pytest.mark.parametrize("N,D,d", [(6, 8, 2), (5, 9, 3)])
@pytest.mark.parametrize("side", ["left", "right"])
def test_orthogonalize_without_normalising(N, D, d, side, rng):
    # leftorthogonalizeQR! / rightorthogonalizeQR! (orthonormalisation.jl:43-56,
    # 77-90): the gauge moves alone, on a non-canonical state of norm != 1.  The
    # state is unchanged, the returned value is its norm^2, every site but the
    # last (first) is an isometry, and the sites equal the oracle's.
    print("case", N, D, d, side, flush=True)
    a, _ = O.mpsrandom(N, D, d, rng)
    a = [1.3 * s + 0.1 * crandn(rng, s.shape) for s in a]
    dense = Dn.mps_to_dense(a)
    ref = copy_sites(a)
    fo = O.leftorthogonalizeQR_ if side == "left" else O.rightorthogonalizeQR_
    fm = M.leftorthogonalizeQR_ if side == "left" else M.rightorthogonalizeQR_
    n_ref = fo(ref)
    g = dev_state(a)
    n_dev = fm(g)
    got = g.to_host()
    assert abs(n_dev - n_ref) < TOL * n_ref
    assert abs(n_dev - np.vdot(dense, dense).real) < TOL * n_dev
    assert rel(Dn.mps_to_dense(got), dense) < TOL
    for j in range(N):
        assert rel(got[j], ref[j]) < 1e-9, j
    for B in (got[:-1] if side == "left" else got[1:]):
        if side == "left":
            Q = B.reshape(-1, B.shape[2], order="F")
            assert np.abs(Q.conj().T @ Q - np.eye(Q.shape[1])).max() < 1e-12
        else:
            Q = B.reshape(B.shape[0], -1, order="F")
            assert np.abs(Q @ Q.conj().T - np.eye(Q.shape[0])).max() < 1e-12


def test_gauge_moves(rng):
    N, D, d = 6, 8, 2
    a, dims = O.mpsrandom(N, D, d, rng)
    a[-1] = a[-1] * 1.7
    # scramble the gauge: non-canonical input
    a = [s + 0.1 * crandn(rng, s.shape) for s in a]
    g = dev_state(a)
    no_ref = O.leftorthonormalizationQR_(copy_sites(a))
    assert abs(M.leftorthonormalizationQR_(g) - no_ref) < TOL * no_ref
    ref = copy_sites(a)
    O.leftorthonormalizationQR_(ref)
    assert rel(Dn.mps_to_dense(g.to_host()), Dn.mps_to_dense(ref)) < TOL
    g2 = dev_state(a)
    no_r = O.rightorthonormalizationQR_(copy_sites(a))
    assert abs(M.rightorthonormalizationQR_(g2) - no_r) < TOL * no_r
    g3 = dev_state(a)
    assert abs(M.norm_lo(g3) - O.norm_lo(a)) < TOL
    assert abs(M.normalize_lo_(g3) - O.normalize_lo_(copy_sites(a))) < TOL


def test_product_state_constructor():
    N, D, d = 9, 8, 3
    st = np.array([0.6, 0.8j, 0.0])
    ref, rdims = O.mpsproductstate(N, D, d, st)
    g, gdims = M.mpsproductstate(N, D, d, st)
    assert list(gdims) == list(rdims)
    got = g.to_host()
    for j in range(N):
        assert rel(got[j], ref[j]) < 1e-14


def test_compress_var_and_sum_var(rng):
    N, D, d = 7, 8, 2
    target, dims = O.mpsrandom(N, D, d, rng)
    guess, _ = O.mpsrandom(N, D, d, rng)
    A0 = copy_sites(guess)
    env = O.build_env(dims, dims)
    O.compress_var_(A0, target, env, 2)
    gA0, gt = dev_state(guess), dev_state(target)
    M.compress_var_(gA0, gt, M.build_env(dims, dims), 2)
    assert rel(Dn.mps_to_dense(gA0.to_host()), Dn.mps_to_dense(A0)) < TOL

    vecs = [O.mpsrandom(N, D, d, rng)[0] for _ in range(3)]
    coef = [0.3, -0.2 + 0.1j, 1.1]
    ref = O.compress_sum_var(guess, coef, vecs, 2)
    got = M.compress_sum_var(dev_state(guess), coef, [dev_state(v) for v in vecs], 2)
    assert rel(Dn.mps_to_dense(got.to_host()), Dn.mps_to_dense(ref)) < TOL


@pytest.mark.parametrize("N,D0,D,d", [(8, 4, 12, 2), (6, 3, 9, 3)])
def test_compress_var_to_smaller_bonds(N, D0, D, d, rng):
    # the fit compress_var! is for: a guess of smaller bond dimension fitted
    # to a larger target; the environment pairs the two dims vectors
    print("case", N, D0, D, d, flush=True)
    target, dims = O.mpsrandom(N, D, d, rng)
    guess, dims0 = O.mpsrandom(N, D0, d, rng)
    assert max(dims0) < max(dims)
    A0 = copy_sites(guess)
    O.compress_var_(A0, target, O.build_env(dims0, dims), 3)
    gA0, gt = dev_state(guess), dev_state(target)
    M.compress_var_(gA0, gt, M.build_env(dims0, dims), 3)
    assert list(gA0.dims) == list(dims0)
    got = gA0.to_host()
    assert rel(Dn.mps_to_dense(got), Dn.mps_to_dense(A0)) < TOL
    # the fit raised the overlap with the target above the guess's
    t = Dn.mps_to_dense(target)
    ov = lambda s: abs(np.vdot(Dn.mps_to_dense(s), t)) / np.linalg.norm(Dn.mps_to_dense(s))
    assert ov(got) > ov(guess)
    assert abs(M.scal_prod(gA0, gt) - np.vdot(Dn.mps_to_dense(A0), t)) < TOL * np.linalg.norm(t) * 10


def test_compress_var_apply_H_full_sweeps(rng):
    N, D, d = 6, 8, 2
    psi, dims = O.mpsrandom(N, D, d, rng)
    H = random_mpo(rng, N, d, 2)
    wd = [1] + [2] * (N - 1) + [1]
    A0 = copy_sites(psi)
    O.compress_var_apply_H_(A0, O.build_env_op(dims, dims, wd), psi, psi, H, 2)
    g = dev_state(psi)
    gA0 = dev_state(psi)
    M.compress_var_apply_H_(gA0, M.build_env(dims, dims, wd), g, g, M.MPO(H), 2)
    assert rel(Dn.mps_to_dense(gA0.to_host()), Dn.mps_to_dense(A0)) < TOL


def test_arnoldi_expm(rng):
    # real MPO and real state: exp(A) psi by Krylov on the MPS manifold
    N, D, d, numv = 5, 4, 2, 4
    psi, dims = O.mpsrandom(N, D, d, rng)
    psi = [s.real.astype(complex) for s in psi]
    O.leftorthonormalizationQR_(psi)
    H = [np.real(h).astype(complex) * 0.3 for h in random_mpo(rng, N, d, 2)]
    wd = [1] + [2] * (N - 1) + [1]
    rhov = [copy_sites(psi) for _ in range(numv)]
    envop = O.build_env_op(dims, dims, wd)
    cref = O.Arnoldi_expm_(H, rhov, numv, None, envop)
    grhov = [dev_state(psi) for _ in range(numv)]
    cgot = M.Arnoldi_expm_(M.MPO(H), grhov, numv, None, M.build_env(dims, dims, wd))
    assert np.abs(cgot - cref).max() < 1e-9
    assert rel(Dn.mps_to_dense(grhov[0].to_host()), Dn.mps_to_dense(rhov[0])) < 1e-9


def _phase_aligned_rel(a, b):
    """Relative difference of two dense states up to one global phase."""
    ov = np.vdot(b, a)
    ph = ov / abs(ov) if abs(ov) > 0 else 1.0
    return rel(a, b * ph)


@pytest.mark.parametrize("N,D,d,dmax,tol", [(6, 8, 2, 64, 1e-15), (6, 8, 2, 2, 1e-15), (8, 16, 3, 5, 1e-15),
                                            (7, 12, 2, 64, 1e-6), (20, 32, 2, 16, 1e-15)])
def test_compression_svd(rng, N, D, d, dmax, tol):
    # compressionSVD! (compression.jl:1280-1312) on a left-canonical random state:
    # same truncation (dims), same spectra (err_trim, ent_ent, err_glob) and the
    # same truncated state up to the global phase the reference drops with r
    print("case", N, D, d, dmax, tol, flush=True)
    a, _ = O.mpsrandom(N, D, d, rng)
    O.leftorthonormalizationQR_(a)
    ref = copy_sites(a)
    eg_r, et_r, ee_r = O.compressionSVD_(ref, dmax, tol)
    g = dev_state(a)
    eg, et, ee = M.compressionSVD_(g, dmax, tol)
    got = g.to_host()
    assert [s.shape for s in got] == [s.shape for s in ref]
    assert abs(eg - eg_r) < 1e-10
    assert np.abs(et - et_r).max() < 1e-10
    assert np.abs(ee - ee_r).max() < 1e-10
    if N <= 8:
        assert _phase_aligned_rel(Dn.mps_to_dense(got), Dn.mps_to_dense(ref)) < 1e-10
    # right-canonical sites: sum_{s,b} B[a,s,b] conj(B[a',s,b]) = I
    for B in got[1:]:
        M2 = B.reshape(B.shape[0], -1, order="F")
        assert np.abs(M2 @ M2.conj().T - np.eye(B.shape[0])).max() < 1e-12


@pytest.mark.parametrize("tol", [1e-10, 1e-6])
def test_compression_svd_of_an_evolved_spin_chain(tol):
    # the spin-chain case's truncation (N = 20, D = 32, cutoff 1e-10) on a
    # state evolved 40 steps from |g...g>, whose spectra decay fast enough
    # for the cutoff, not D, to set the bonds
    from oracle import trajectory as OT
    N, D, dt = 20, 32, 0.01
    print("case", tol, flush=True)
    z = models.positions(N, 20.0, 1)
    H = models.spin_rk4_mpos(z, 0.0, dt)
    jr, jl = models.spin_jump_mpos(z)
    O.set_qr_rank_tol(1e-13)
    try:
        tr = OT.OracleSpinTrajectory(N, D, 1, H, jr, jl, dt)
        for _ in range(40):
            tr.step()
    finally:
        O.set_qr_rank_tol(0.0)
    ref = copy_sites(tr.A)
    g = dev_state(tr.A)
    eg_r, et_r, ee_r = O.compressionSVD_(ref, D, tol)
    assert max(s.shape[2] for s in ref) < D
    eg, et, ee = M.compressionSVD_(g, D, tol)
    got = g.to_host()
    assert [s.shape for s in got] == [s.shape for s in ref]
    assert abs(eg - eg_r) < 1e-12
    assert np.abs(et - et_r).max() < 1e-12
    assert np.abs(ee - ee_r).max() < 1e-10
    assert _phase_aligned_rel(Dn.mps_to_dense(got), Dn.mps_to_dense(ref)) < 1e-10


@pytest.mark.parametrize("graphs", [False, True])
def test_rk4_fp32_operator_terms(graphs):
    # Precision form: the operator contractions run in complex FP32 (3M MFMA,
    # FP32 accumulation) while QR, norms and overlap terms stay FP64.  Against
    # the FP64 oracle the error must sit at the FP32 level: above FP64 round-off
    # (the FP32 kernel really ran) and within 1e-5 of the dense state.
    N, D, dt, steps = 8, 8, 0.01, 4
    print("case fp32", graphs, flush=True)
    z, H, psi, dims = _spin_setup(N, D, dt, "random")
    wop = [1] + [4] * (N - 1) + [1]
    ref, rn = _oracle_rk4(H, psi, dims, dt, steps, wop)
    ctx = M.default_context()
    ctx.set_precision("fp32")
    try:
        g, gn = _device_rk4(H, psi, dims, dt, steps, wop, False, graphs)
    finally:
        ctx.set_precision("fp64")
    err = rel(Dn.mps_to_dense(g.to_host()), Dn.mps_to_dense(ref))
    print("fp32 relative state error", err, flush=True)
    assert 1e-12 < err < 1e-5
    assert np.allclose(gn, rn, rtol=1e-5, atol=0)


def test_environment_slots_after_compress_var_init(rng):
    # mps_env_download_slot: compress_var! with no sweeps only builds its left
    # overlap environments (compression.jl:65-103, the initialisation loop);
    # every slot equals the oracle's, boundaries are 1
    N, D0, D, d = 6, 4, 8, 2
    target, dims = O.mpsrandom(N, D, d, rng)
    guess, dims0 = O.mpsrandom(N, D0, d, rng)
    env_o = O.build_env(dims0, dims)
    O.compress_var_(copy_sites(guess), target, env_o, 0)
    env_d = M.build_env(dims0, dims)
    M.compress_var_(dev_state(guess), dev_state(target), env_d, 0)
    for j in range(N + 1):
        ref = np.asarray(env_o[j])
        got = env_d.slot(j, ref.shape)
        assert rel(got, ref) < TOL, j
    assert env_d.slot(0, (1, 1))[0, 0] == 1.0 and env_d.slot(N, (1, 1))[0, 0] == 1.0


def test_kernel_stats_count_the_algorithmic_gemm_work(rng):
    # the measurement extension: with profiling on, every GEMM launch is timed
    # and its algorithmic flops (8 per complex multiply-add) are summed; for
    # <A|B> that is, per site, L*B (Da_l x d Db_r x Db_l) and the environment
    # update (Da_r x Db_r x Da_l d)
    N, D, d = 7, 8, 3
    a, da = O.mpsrandom(N, D, d, rng)
    b, db = O.mpsrandom(N, 5, d, rng)
    ga, gb = dev_state(a), dev_state(b)
    ctx = M.default_context()
    ref = M.scal_prod(ga, gb)
    ctx.set_profiling(True)
    try:
        ctx.kernel_stats()
        got = M.scal_prod(ga, gb)
        st = ctx.kernel_stats()
    finally:
        ctx.set_profiling(False)
    flops = sum(8.0 * da[j] * d * db[j + 1] * db[j] + 8.0 * da[j + 1] * db[j + 1] * da[j] * d for j in range(N))
    print(st, flops, flush=True)
    assert got == ref
    assert st["gemm_launches"] == 2 * N
    assert st["gemm_flops"] == flops
    assert st["gemm_ms"] > 0
