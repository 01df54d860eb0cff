"""Pins the oracle's quantum-jump trajectory loop against a dense state-vector restatement."""
import numpy as np

from oracle import dense as Dn
from oracle import trajectory as OT
from mpsamd import models


def test_oracle_trajectory_matches_dense_jump_process():
    # N = 3 at full bond dimension: every compression is exact, so the MPS
    # trajectory must follow the dense one (same draws) step for step:
    # psi_coh = RK4(psi), p_coh = ||psi_coh||^2, jumps E_R psi, E_L psi, sigma^-_j psi
    N, D, dt, steps, seed, gam_eg = 3, 4, 0.08, 40, 3, 0.7
    z = models.positions(N, 3.0, seed)
    H = models.spin_rk4_mpos(z, 0.0, dt)
    jr, jl = models.spin_jump_mpos(z)
    tr = OT.OracleSpinTrajectory(N, D, seed, H, jr, jl, dt, gam_eg, init=(0.0, 1.0))
    Lm = Dn.mpo_to_dense(H[1])                 # -i H_eff
    ER, EL = Dn.mpo_to_dense(jr), Dn.mpo_to_dense(jl)
    psi = Dn.mps_to_dense(tr.A)
    rng = np.random.default_rng(seed)
    kinds = []
    for i in range(steps):
        k1 = Lm @ psi
        k2 = Lm @ (psi + dt / 2 * k1)
        k3 = Lm @ (psi + dt / 2 * k2)
        k4 = Lm @ (psi + dt * k3)
        coh = psi + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        p = np.vdot(coh, coh).real
        p_mps = tr.step()
        assert abs(p_mps - p) < 1e-12
        r1, r2 = rng.random(2)
        if r1 < p:
            psi = coh / np.sqrt(p)
            continue
        pr, pl = np.linalg.norm(ER @ psi) ** 2, np.linalg.norm(EL @ psi) ** 2
        epop = [np.vdot(psi, Dn.site_op(N, 2, j, models.HEE2) @ psi).real for j in range(N)]
        acc = np.concatenate([[0.0], np.cumsum(dt * gam_eg * np.array(epop))])
        acc = np.concatenate([acc, [acc[-1] + dt * pr, acc[-1] + dt * pr + dt * pl]])
        acc /= acc[-1]
        if r2 < acc[N]:
            j = int(np.argmax(r2 < acc[1:N + 1]))
            v = Dn.site_op(N, 2, j, models.HGE2) @ psi
            kinds.append("eg")
        elif r2 < acc[N + 1]:
            v = ER @ psi
            kinds.append("r")
        else:
            v = EL @ psi
            kinds.append("l")
        psi = v / np.linalg.norm(v)
    assert [k for _, k, _ in tr.jumps] == kinds and len(kinds) >= 2
    got = Dn.mps_to_dense(tr.A)
    ov = np.vdot(got, psi)
    assert abs(abs(ov) - 1) < 1e-10
    sz = [np.vdot(psi, Dn.site_op(N, 2, j, 2 * models.HEE2 - np.eye(2)) @ psi).real for j in range(N)]
    assert np.abs(tr.sigma_z() - np.array(sz)).max() < 1e-10
