"""The oracle's steady_state_search! (steadystate.jl:6-86) pinned by the
dense Lindbladian L of a three-atom Rydberg ladder at full bond dimension:

* the search's last local eigenvalue ae is <rho|L^dag L|rho> / <rho|rho>,
  i.e. ||L rho||^2 / ||rho||^2 computed densely (this pins the two-operator
  environments and the local operator, MatrixProductStates.jl:258-312);
* the result is the null vector of L (the unique steady state) within the
  bound the residual sets, sin(angle) <= ||L rho|| / (sigma_2 ||rho||), with
  sigma_2 the second-smallest singular value of L."""
import numpy as np

from mpsamd import models
from oracle import dense as Dn
from oracle import mps_oracle as O


def test_steady_state_search_finds_the_dense_null_vector():
    N = 3
    p = models.RydbergParams(n_exp=2)
    z = models.positions(N, 3.0, 6)
    L = models.rydberg_lindblad_mpo(p, z, 0.0, 1.0)
    Lc = O.conj_mpo(L)
    At, dims = O.mpsrandom(N, 9, 9, np.random.default_rng(3))     # full bond dimension
    ae, sw = O.steady_state_search_(At, L, Lc, 8, 2.0)
    Ld = Dn.mpo_to_dense(L)
    _, S, Vh = np.linalg.svd(Ld)
    ss, sigma2 = Vh[-1].conj(), S[-2]
    x = Dn.mps_to_dense(At)
    r = np.linalg.norm(Ld @ x) / np.linalg.norm(x)
    cos = abs(np.vdot(ss, x)) / np.linalg.norm(x)
    sin = np.sqrt(max(0.0, 1.0 - cos * cos))
    print("ae", ae, "sweeps", sw, "||L x||/||x||", r, "sin", sin, "bound", r / sigma2, flush=True)
    assert abs(ae - r * r) < 1e-12
    assert r < 1e-5 and S[-1] < 1e-12 * S[0]
    assert sin <= 1.01 * r / sigma2 + 1e-12
