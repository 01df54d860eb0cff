"""steady_state_search! (steadystate.jl:6-86) on the device: the null vector
of a three-atom Rydberg Lindbladian at full bond dimension within the bound
its residual sets (as the oracle's pin), and the same search at truncated bond
dimension against the oracle (state fidelity and final local eigenvalue)."""
import numpy as np
import pytest

import mpsamd as M
from mpsamd import models
from oracle import dense as Dn
from oracle import mps_oracle as O
from helpers import copy_sites

pytestmark = pytest.mark.gpu


def test_steady_state_full_bond_null_vector():
    N = 3
    p = models.RydbergParams(n_exp=2)
    z = models.positions(N, 3.0, 6)
    L = models.rydberg_lindblad_mpo(p, z, 0.0, 1.0)
    At, dims = O.mpsrandom(N, 9, 9, np.random.default_rng(3))
    g = M.MPS.from_host(At)
    ae, sw = M.steady_state_search_(g, L, O.conj_mpo(L), 8, 2.0, 1e-13)
    Ld = Dn.mpo_to_dense(L)
    _, S, Vh = np.linalg.svd(Ld)
    x = Dn.mps_to_dense(g.to_host())
    r = np.linalg.norm(Ld @ x) / np.linalg.norm(x)
    cos = abs(np.vdot(Vh[-1].conj(), x)) / np.linalg.norm(x)
    sin = np.sqrt(max(0.0, 1.0 - cos * cos))
    print("device ae", ae, "sweeps", sw, "residual", r, "sin", sin, "bound", r / S[-2], flush=True)
    assert r < 1e-5 and abs(ae - r * r) < 1e-12
    assert sin <= 1.01 * r / S[-2] + 1e-12


def test_steady_state_truncated_against_oracle():
    N, D = 5, 12
    p = models.RydbergParams(n_exp=2)
    z = models.positions(N, 5.0, 2)
    L = models.rydberg_lindblad_mpo(p, z, 0.0, 1.0)
    Lc = O.conj_mpo(L)
    At, dims = O.mpsrandom(N, D, 9, np.random.default_rng(5))
    ref = copy_sites(At)
    ae_o, sw_o = O.steady_state_search_(ref, L, Lc, 3, 2.0)
    g = M.MPS.from_host(At)
    ae, sw = M.steady_state_search_(g, L, Lc, 3, 2.0, 1e-13)
    got = g.to_host()
    ab = O.scal_prod(got, ref)
    infid = 1.0 - abs(ab) ** 2 / (abs(O.scal_prod(got, got)) * abs(O.scal_prod(ref, ref)))
    print("ae", ae, ae_o, "sweeps", sw, sw_o, "1 - fidelity", infid, flush=True)
    assert sw == sw_o
    assert abs(ae - ae_o) < 1e-10 * max(1.0, abs(ae_o))
    assert infid < 1e-8
