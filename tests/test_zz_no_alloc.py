"""rcg_residuals_noalloc: the residual norms of rcg_residuals computed in the
library's fixed device workspace, with no allocation.

Pins: through the C ABI, two calls in a row (the second on a smaller operator
and fewer columns, so it reuses the workspace the first one left) give the
residuals of rcg_residuals to a relative 1e-14, for a real and a complex
operator; both agree with scipy's r_c = ||theta_c v_c - A v_c||; a request
beyond the workspace is refused with RCG_ERR_INVALID and changes nothing.
"""
import ctypes as C

import numpy as np
import pytest

from rayleighchebyshevgpgpu_amd import capi
from rayleighchebyshevgpgpu_amd import operators as ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rcg():
    from rayleighchebyshevgpgpu_amd import solver
    return solver


def residuals_both(rcg, A, cols, seed, cplx):
    op = rcg.CsrOperator(A)
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((A.n, cols))
    if cplx:
        V = V + 1j * rng.standard_normal((A.n, cols))
    bv = rcg.Block.from_host(V)
    bw1 = rcg.Block(A.n, cols, dtype=bv.dtype)
    bw2 = rcg.Block(A.n, cols, dtype=bv.dtype)
    th = rng.standard_normal(cols)
    r_alloc = np.zeros(cols)
    r_none = np.full(cols, np.nan)
    L = capi.lib()
    capi.check(L.rcg_residuals(op.h, bv.h, bw1.h, cols, th.ctypes.data_as(C.c_void_p),
                               r_alloc.ctypes.data_as(C.c_void_p)))
    capi.check(L.rcg_residuals_noalloc(op.h, bv.h, bw2.h, cols, th.ctypes.data_as(C.c_void_p),
                                       r_none.ctypes.data_as(C.c_void_p)))
    rref = np.linalg.norm(th * V - A.to_scipy() @ V, axis=0)
    # W = A V comes back in the caller's block, as from rcg_residuals
    assert np.array_equal(bw1.download(), bw2.download())
    return r_alloc, r_none, rref


@pytest.mark.parametrize("cplx", [False, True])
def test_noalloc_matches_residuals_twice(rcg, cplx):
    if cplx:
        big = ops.hofstadter_2d(48, phi=1.0 / 8, W=1.0, seed=2)
        small = ops.hofstadter_2d(16, phi=1.0 / 4, W=1.0, seed=3)
    else:
        big = ops.laplacian_3d(20, bc="periodic", potential=ops.anderson_potential(8000, 4.0, 1))
        small = ops.laplacian_2d(30)
    for call, (A, cols) in enumerate([(big, 40), (small, 7)]):
        print(f"case: no-alloc residuals call {call}, n={A.n}, cols={cols}, complex={cplx}",
              flush=True)
        r_alloc, r_none, rref = residuals_both(rcg, A, cols, 10 + call, cplx)
        assert np.all(np.isfinite(r_none))
        assert np.max(np.abs(r_none - r_alloc) / r_alloc) <= 1e-14, call
        assert np.max(np.abs(r_none - rref) / rref) < 1e-12, call


def test_noalloc_refuses_beyond_workspace(rcg):
    # 2^21 doubles: 3 per column plus one per column and 512 rows; n = 2^20
    # rows give 2048 partial rows, so 1100 columns need 2051 x 1100 > 2^21
    n, cols = 1 << 20, 1100
    print(f"case: no-alloc residuals beyond the workspace, n={n}, cols={cols}", flush=True)
    A = ops.laplacian_2d(1024)
    op = rcg.CsrOperator(A)
    bv = rcg.Block(n, cols)
    bw = rcg.Block(n, cols)
    th = np.zeros(cols)
    res = np.full(cols, -1.0)
    code = capi.lib().rcg_residuals_noalloc(op.h, bv.h, bw.h, cols, th.ctypes.data_as(C.c_void_p),
                                            res.ctypes.data_as(C.c_void_p))
    assert code == 1
    assert "workspace" in capi.last_error()
    assert np.all(res == -1.0)
