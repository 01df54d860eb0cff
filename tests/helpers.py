"""Shared builders for the parity tests: identical seeded inputs for the oracle and the device."""
import numpy as np

from oracle import mps_oracle as O
from oracle import dense as Dn


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def crandn(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * np.sqrt(0.5)


def random_mpo(rng, N, d, w):
    sites = []
    for j in range(N):
        wl = 1 if j == 0 else w
        wr = 1 if j == N - 1 else w
        sites.append(crandn(rng, (wl, d, wr, d)) / np.sqrt(w * d))
    return sites


def dense_state(mps_sites):
    return Dn.mps_to_dense(mps_sites)


def copy_sites(s):
    return [a.copy() for a in s]


def oracle_random_state(N, D, d, seed):
    mps, dims = O.mpsrandom(N, D, d, np.random.default_rng(seed))
    return mps, dims
