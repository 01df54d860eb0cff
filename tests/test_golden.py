"""The oracle against the committed golden fixtures (tests/golden/, written by
make_golden.py from closed forms, dense linear algebra and LAPACK)."""
import os

import numpy as np

from oracle import mps_oracle as O
from oracle import dense as Dn
from mpsamd import models
from helpers import rel

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def _one_excited(n):
    return np.array([0, 1], dtype=complex) if n == 1 else np.array([1, 0], dtype=complex)


def _rk4_norms(H, psi, dims, dt, steps, w):
    envop = O.build_env_op(dims, dims, [1] + [w] * (len(psi) - 1) + [1])
    e = [O.build_env(dims, dims) for _ in range(3)]
    p1, p2, p3 = ([a.copy() for a in psi] for _ in range(3))
    norms = []
    for _ in range(steps):
        O.RK4stp_apply_H_half_(dt, H, psi, p1, p2, p3, *e, envop)
        norms.append(O.normalize_lo_(psi))
    return np.array(norms)


def test_single_excitation_prefix():
    # the first 24 sites of the N = 100 fixture's recurrence at N = 24 would be a
    # different chain, so the oracle runs the full N = 100 chain at D = 2 for 4 steps
    f = _load("single_excitation_N100.npz")
    z, dt = f["z"], float(f["dt"])
    N = len(z)
    H = models.spin_rk4_mpos(z, 0.0, dt, Omega=0.0)
    psi, dims = O.mpsproductstate(N, 2, 2, _one_excited)
    norms = _rk4_norms(H, psi, dims, dt, 4, 4)
    assert np.abs(norms / f["norms"][:4] - 1).max() < 1e-12


def test_rk4_dense_fixture():
    f = _load("rk4_dense_N6.npz")
    z, dt, steps = f["z"], float(f["dt"]), int(f["steps"])
    H = models.spin_rk4_mpos(z, 0.0, dt)
    psi, dims = O.mpsproductstate(6, 8, 2, np.array([1, 0], dtype=complex))
    norms = _rk4_norms(H, psi, dims, dt, steps, 4)
    assert np.abs(norms / f["norms"] - 1).max() < 1e-12
    assert rel(Dn.mps_to_dense(psi), f["states"][-1]) < 1e-11


def test_qr_fixture():
    f = _load("qr_lapack.npz")
    assert np.abs(O.qrq(f["M"]) - f["Q"]).max() < 1e-13


def dense_to_mps(v, N, d=2):
    """Left-canonical exact MPS of a dense vector (site 1 fastest) by successive QR."""
    sites, r, Dl = [], v.reshape(1, -1), 1
    for n in range(N - 1):
        m = r.reshape(Dl * d, -1, order="F")
        q, rr = np.linalg.qr(m)
        sites.append(q.reshape(Dl, d, q.shape[1], order="F"))
        Dl, r = q.shape[1], rr
    sites.append(r.reshape(Dl, d, 1, order="F"))
    return sites


def test_svd_fixture():
    f = _load("svd_dense_N6.npz")
    mps = dense_to_mps(f["psi"], 6)
    assert rel(Dn.mps_to_dense(mps), f["psi"]) < 1e-12
    eg, et, ee = O.compressionSVD_(mps, 64, 1e-15)
    for cut in range(1, 6):
        s = f[f"s{cut}"]
        # ent_ent[i] belongs to the cut left of site i (0-based)
        assert abs(ee[cut] - (-2 * np.sum(s ** 2 * np.log(s)))) < 1e-10
    assert abs(eg) < 1e-12
