"""rcg_cheb_filter_batch: the filters of several CSR operators in shared
launches (one per step and launch class), as the server runs them.

Pins: each operator's result equals rcg_cheb_filter's on that operator bit for
bit (the batched step computes every row as the single-operator step does)
and the oracle filter (LanczosCpolyOperatorLM.h:192-304) to 1e-12, for real
and complex operators of different sizes, widths (launch classes) and
degrees, odd and even, including degree 1 and an empty column count."""
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


def arr(ptype, items):
    return (ptype * len(items))(*items)


def test_batched_filter_equals_single_and_oracle(rcg, orc):
    print("case: batched filter of 9 operators vs rcg_cheb_filter and the oracle", flush=True)
    rng = np.random.default_rng(5)
    specs = [  # operator, columns, degree, lambda_max, shift
        (ops.laplacian_3d(12, bc="periodic", potential=ops.anderson_potential(1728, 4.0, 1)), 24, 7, 14.5, 1.5),
        (ops.laplacian_2d(40, potential=ops.anderson_potential(1600, 2.0, 3)), 12, 12, 9.0, 0.2),
        (ops.laplacian_2d(23), 3, 1, 8.1, 0.1),
        (ops.laplacian_3d(9, bc="dirichlet"), 48, 20, 12.2, -0.3),
        (ops.hofstadter_2d(16, phi=1.0 / 8, W=1.0, seed=2), 10, 9, 9.0, 4.1),
        (ops.laplacian_2d(30), 0, 5, 8.1, 0.0),
        # lanes per row that do not divide the workgroup: 3, 48 (real), 24 (complex)
        (ops.laplacian_2d(37, potential=ops.anderson_potential(1369, 1.0, 4)), 5, 6, 8.6, 0.3),
        (ops.laplacian_3d(7, bc="dirichlet"), 80, 4, 12.1, 0.0),
        (ops.hofstadter_2d(12, phi=1.0 / 4, W=0.5, seed=3), 20, 5, 8.6, 4.0),
    ]
    held, Xs, Ts, refs = [], [], [], []
    for A, cols, m, lmax, shift in specs:
        op = rcg.CsrOperator(A)
        w = max(cols, 1)
        X = rng.uniform(-1, 1, (A.n, w))
        if A.is_complex:
            X = X + 1j * rng.uniform(-1, 1, (A.n, w))
        bx = rcg.Block.from_host(X)
        bt = rcg.Block(A.n, w, dtype=bx.dtype)
        bs, bst = rcg.Block.from_host(X), rcg.Block(A.n, w, dtype=bx.dtype)
        if cols:
            capi.check(capi.lib().rcg_cheb_filter(op.h, bs.h, bst.h, cols, m, 1, lmax, shift))
        held.append((A, op, cols, m, lmax, shift, X, bs))
        Xs.append(bx)
        Ts.append(bt)
    n = len(specs)
    capi.check(capi.lib().rcg_cheb_filter_batch(
        n, arr(C.c_void_p, [h[1].h.value for h in held]), arr(C.c_void_p, [b.h.value for b in Xs]),
        arr(C.c_void_p, [b.h.value for b in Ts]), arr(C.c_int64, [h[2] for h in held]),
        arr(C.c_int64, [h[3] for h in held]), arr(C.c_double, [h[4] for h in held]),
        arr(C.c_double, [h[5] for h in held])))
    for (A, op, cols, m, lmax, shift, X, bs), bx in zip(held, Xs):
        got = bx.download()
        if cols == 0:
            assert np.array_equal(got, X)
            continue
        single = bs.download()
        assert np.array_equal(got[:, :cols], single[:, :cols]), (A.n, cols, m)
        ref = orc.filter_csr(A, X[:, :cols], m, lmax, shift)
        assert np.max(np.abs(got[:, :cols] - ref)) <= 1e-12 * np.max(np.abs(ref)), (A.n, cols, m)


def test_batched_filter_rejects_bad_input(rcg):
    A = ops.laplacian_2d(10)
    op = rcg.CsrOperator(A)
    D = rcg.DenseOperator(np.eye(A.n))
    bx, bt = rcg.Block(A.n, 4), rcg.Block(A.n, 4)
    L = capi.lib()
    one = lambda t, v: arr(t, [v])  # noqa: E731
    # a dense operator, and a degree of 0
    assert L.rcg_cheb_filter_batch(1, one(C.c_void_p, D.h.value), one(C.c_void_p, bx.h.value),
                                   one(C.c_void_p, bt.h.value), one(C.c_int64, 4),
                                   one(C.c_int64, 3), one(C.c_double, 8.0),
                                   one(C.c_double, 0.0)) == 1
    assert L.rcg_cheb_filter_batch(1, one(C.c_void_p, op.h.value), one(C.c_void_p, bx.h.value),
                                   one(C.c_void_p, bt.h.value), one(C.c_int64, 4),
                                   one(C.c_int64, 0), one(C.c_double, 8.0),
                                   one(C.c_double, 0.0)) == 1
