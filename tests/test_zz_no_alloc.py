"""mps_scal_prod_no_conj_noalloc (the entry point that allocates nothing)
against mps_scal_prod_no_conj, through the C ABI: two calls in a row, the
second with smaller states, then inputs too large for the fixed buffer."""
import ctypes as C

import numpy as np
import pytest

import mpsamd as M
from mpsamd import _lib
from oracle import mps_oracle as O

pytestmark = pytest.mark.gpu


def _call(name, a, b):
    out = np.zeros(2)
    rc = getattr(_lib.load(), name)(a.h, b.h, out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, complex(out[0], out[1])


@pytest.mark.parametrize("N,D,d", [(8, 64, 9), (5, 12, 9)])
def test_noalloc_matches_scal_prod_no_conj_twice(N, D, d):
    ctx = M.Context(0)
    rng = np.random.default_rng(N * 1000 + D)
    for k, (n, dm) in enumerate(((N, D), (max(2, N - 3), max(2, D // 2)))):   # second call smaller
        print("case no-alloc scal_prod_no_conj N", n, "D", dm, "d", d, flush=True)
        a, _ = O.mpsrandom(n, dm, d, rng)
        b, _ = O.mpsrandom(n, dm + 3, d, rng)
        ga, gb = M.MPS.from_host(a, ctx), M.MPS.from_host(b, ctx)
        rc0, ref = _call("mps_scal_prod_no_conj", ga, gb)
        rc1, got = _call("mps_scal_prod_no_conj_noalloc", ga, gb)
        assert rc0 == 0 and rc1 == 0
        assert abs(got - ref) <= 1e-10 * abs(ref), (k, got, ref)
        assert abs(ref - O.scal_prod_no_conj(a, b)) <= 1e-10 * abs(ref)


def test_noalloc_rejects_inputs_beyond_the_buffer():
    ctx = M.Context(0)
    N, D, d = 3, 512, 9     # per-site product 512 x 9 x 512 complex = 36 MiB > 24 MiB
    dims = [1, D, D, 1]
    ga, gb = M.MPS(N, d, dims, ctx), M.MPS(N, d, dims, ctx)
    rc, _ = _call("mps_scal_prod_no_conj_noalloc", ga, gb)
    assert rc == _lib.MPS_E_ARG
    assert b"fixed device buffer" in _lib.load().mps_last_error()
