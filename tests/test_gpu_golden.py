"""The device against the committed golden fixtures, including the full N = 100
single-excitation chain (a closed form that holds at any bond dimension >= 2)."""
import os

import numpy as np
import pytest

import mpsamd as M
from mpsamd import models
from oracle import dense as Dn
from helpers import rel
from test_golden import dense_to_mps

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def _device_rk4(H, psi, dims, dt, steps, w):
    gH = [M.MPO(h) for h in H]
    r = [psi.similar() for _ in range(3)]
    for s in r:
        s.copy_from(psi)
    e = [M.build_env(dims, dims) for _ in range(3)]
    eop = M.build_env(dims, dims, [1] + [w] * (len(dims) - 2) + [1])
    norms = []
    for _ in range(steps):
        M.RK4stp_apply_H_half_(dt, gH, psi, *r, *e, eop)
        norms.append(M.normalize_lo_(psi))
    return np.array(norms)


@pytest.mark.parametrize("D,tol", [(2, 0.0), (16, 1e-13)])
def test_single_excitation_N100(D, tol):
    # D = 16 is 8x the state's true rank: with the plain zlarfg rule the QR
    # completions come from rounding noise, and at this size the one-site RK4
    # amplifies them after ~16 steps (the oracle, run from the device's state,
    # reproduces the same step); the rank cutoff makes the completion exact
    f = _load("single_excitation_N100.npz")
    z, dt, steps = f["z"], float(f["dt"]), int(f["steps"])
    N = len(z)
    print("case single excitation N=100 D", D, tol, flush=True)
    ctx = M.default_context()
    ctx.set_qr_rank_tol(tol)
    try:
        H = models.spin_rk4_mpos(z, 0.0, dt, Omega=0.0)
        psi, dims = M.mpsproductstate(N, D, 2,
                                      lambda n: np.array([0, 1], complex) if n == 1 else np.array([1, 0], complex))
        norms = _device_rk4(H, psi, dims, dt, steps, 4)
        n_e = M.measure_excitations_(None, psi, models.HEE2)
    finally:
        ctx.set_qr_rank_tol(0.0)
    assert np.abs(norms / f["norms"] - 1).max() < 1e-10
    assert np.abs(n_e - f["n_e"]).max() < 1e-10


def test_rk4_dense_fixture():
    f = _load("rk4_dense_N6.npz")
    z, dt, steps = f["z"], float(f["dt"]), int(f["steps"])
    H = models.spin_rk4_mpos(z, 0.0, dt)
    psi, dims = M.mpsproductstate(6, 8, 2, np.array([1, 0], dtype=complex))
    norms = _device_rk4(H, psi, dims, dt, steps, 4)
    assert np.abs(norms / f["norms"] - 1).max() < 1e-12
    assert rel(Dn.mps_to_dense(psi.to_host()), f["states"][-1]) < 1e-11


def test_qr_fixture():
    import devbuf as dev
    f = _load("qr_lapack.npz")
    assert np.abs(dev.qrq(f["M"]) - f["Q"]).max() < 1e-12


def test_svd_fixture():
    f = _load("svd_dense_N6.npz")
    g = M.MPS.from_host(dense_to_mps(f["psi"], 6))
    eg, et, ee = M.compressionSVD_(g, 64, 1e-15)
    for cut in range(1, 6):
        s = f[f"s{cut}"]
        assert abs(ee[cut] - (-2 * np.sum(s ** 2 * np.log(s)))) < 1e-10
    assert abs(eg) < 1e-12
