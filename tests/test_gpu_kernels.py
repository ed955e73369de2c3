"""HIP kernels of librcg.so against the CPU oracle and numpy (FP64).

Each test names its case before touching the device.  Tolerances: FP64
kernels agree with the oracle / numpy to 1e-12 relative (they differ only in
summation order and fused multiply-adds).
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


def anderson3d(nx, W=4.0, seed=1):
    pot = ops.anderson_potential(nx ** 3, W, seed)
    return ops.laplacian_3d(nx, bc="periodic", potential=pot)


def test_block_roundtrip(rcg):
    print("case: block upload/download 777x13", flush=True)
    X = np.random.default_rng(0).standard_normal((777, 13))
    b = rcg.Block.from_host(X)
    assert b.ld % 4 == 0 and b.ld >= 13
    assert np.array_equal(b.download(), X)


def test_randomize_bit_exact(rcg):
    print("case: randomize cols 2..6 of 1000x12, seed 3, counter 100", flush=True)
    n = 1000
    b = rcg.Block(n, 12)
    c = b.randomize(2, 5, seed=3, counter=100)
    assert c == 100 + 5 * n
    X = b.download()
    for j in range(5):
        assert np.array_equal(X[:, 2 + j], ops.sym_draws(3, 100 + j * n, n))
    assert np.all(X[:, :2] == 0) and np.all(X[:, 7:] == 0)


@pytest.mark.parametrize("cols", [1, 3, 12, 64, 100])
def test_csr_apply_vs_scipy(rcg, cols):
    print(f"case: csr apply 10^3 Anderson, cols={cols}", flush=True)
    A = anderson3d(10)
    op = rcg.CsrOperator(A)
    rng = np.random.default_rng(cols)
    X = rng.standard_normal((A.n, cols))
    Y0 = rng.standard_normal((A.n, cols))
    bx, by = rcg.Block.from_host(X), rcg.Block.from_host(Y0)
    op.apply(bx, by, 2.0, 0.5)
    ref = 2.0 * (A.to_scipy() @ X) + 0.5 * Y0
    got = by.download()
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("m,cols", [(1, 5), (2, 16), (7, 64), (30, 48), (31, 8)])
def test_cheb_filter_vs_oracle(rcg, orc, m, cols):
    print(f"case: cheb filter degree {m}, cols {cols}", flush=True)
    A = anderson3d(12)
    op = rcg.CsrOperator(A)
    X = np.random.default_rng(m).uniform(-1, 1, (A.n, cols))
    lmax, shift = 14.5, 1.7
    ref = orc.filter_csr(A, X, m, lmax, shift)
    bx, bt = rcg.Block.from_host(X), rcg.Block(A.n, cols)
    capi.check(capi.lib().rcg_cheb_filter(op.h, bx.h, bt.h, cols, m, 1, lmax, shift))
    got = bx.download()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_cheb_filter_repetition(rcg, orc):
    print("case: cheb filter degree 5 repetition 2", flush=True)
    A = anderson3d(8)
    op = rcg.CsrOperator(A)
    X = np.random.default_rng(9).uniform(-1, 1, (A.n, 6))
    ref = orc.filter_csr(A, X, 5, 13.0, 2.0, rep=2)
    bx, bt = rcg.Block.from_host(X), rcg.Block(A.n, 6)
    capi.check(capi.lib().rcg_cheb_filter(op.h, bx.h, bt.h, 6, 5, 2, 13.0, 2.0))
    got = bx.download()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("n,cols", [(5000, 40), (20000, 64), (3000, 100), (6000, 200), (9000, 256)])
def test_orthonormalize(rcg, n, cols):
    print(f"case: CholQR2 {n}x{cols}", flush=True)
    X = np.random.default_rng(n).standard_normal((n, cols))
    bx, bt = rcg.Block.from_host(X), rcg.Block(n, cols)
    capi.check(capi.lib().rcg_orthonormalize(bx.h, bt.h, cols))
    Q = bx.download()
    assert np.max(np.abs(Q.T @ Q - np.eye(cols))) < 1e-13
    R = Q.T @ X
    assert np.max(np.abs(np.tril(R, -1))) < 1e-10 * np.max(np.abs(R))
    assert np.all(np.diag(R) > 0)
    # same Q as modified Gram-Schmidt (the reference's orthogonalize)
    Qm, Rm = np.linalg.qr(X)
    Qm = Qm * np.sign(np.diag(Rm))
    assert np.max(np.abs(Q - Qm)) < 1e-10


def test_orthonormalize_ill_conditioned(rcg):
    print("case: CholQR on a block with condition 1e12", flush=True)
    rng = np.random.default_rng(5)
    n, b = 4000, 24
    U, _ = np.linalg.qr(rng.standard_normal((n, b)))
    Vt, _ = np.linalg.qr(rng.standard_normal((b, b)))
    X = U @ np.diag(np.logspace(0, -12, b)) @ Vt
    bx, bt = rcg.Block.from_host(X), rcg.Block(n, b)
    capi.check(capi.lib().rcg_orthonormalize(bx.h, bt.h, b))
    Q = bx.download()
    assert np.max(np.abs(Q.T @ Q - np.eye(b))) < 1e-12


@pytest.mark.parametrize("cols,found", [(16, 8), (40, 100), (24, 220)])
def test_deflate(rcg, cols, found):
    print(f"case: deflate {cols} cols against {found} found", flush=True)
    rng = np.random.default_rng(2)
    n = 6000
    Y, _ = np.linalg.qr(rng.standard_normal((n, found)))
    X = rng.standard_normal((n, cols))
    bx, by = rcg.Block.from_host(X), rcg.Block.from_host(Y)
    capi.check(capi.lib().rcg_deflate(bx.h, cols, by.h, found))
    Z = bx.download()
    assert np.max(np.abs(Y.T @ Z)) < 1e-14
    assert np.allclose(np.linalg.norm(Z, axis=0), 1.0, atol=1e-14)
    P = X - Y @ (Y.T @ X)
    P /= np.linalg.norm(P, axis=0)
    assert np.max(np.abs(Z - P)) < 1e-13


@pytest.mark.parametrize("cplx", [False, True])
def test_residuals_on_untuned_wide_block(rcg, cplx):
    """Residuals / apply through the piecewise entry points on a fresh operator at a
    width large enough for the filter autotuner (n * columns >= 2^22): the entry
    points hold no scratch block, so they must run untuned."""
    print(f"case: residuals on a fresh 40^3 operator, 112 columns, complex={cplx}", flush=True)
    if cplx:
        A = ops.hofstadter_2d(256, phi=1 / 64, W=1.0, seed=2)
    else:
        A = anderson3d(40)
    cols = 112
    S = A.to_scipy()
    op = rcg.CsrOperator(A)
    rng = np.random.default_rng(5)
    V = rng.standard_normal((A.n, cols))
    if cplx:
        V = V + 1j * rng.standard_normal((A.n, cols))
    bv = rcg.Block.from_host(V)
    bw = rcg.Block(A.n, cols, dtype=bv.dtype)
    th = rng.standard_normal(cols)
    res = np.zeros(cols)
    capi.check(capi.lib().rcg_residuals(op.h, bv.h, bw.h, cols, th.ctypes.data_as(C.c_void_p),
                                        res.ctypes.data_as(C.c_void_p)))
    rref = np.linalg.norm(th * V - S @ V, axis=0)
    assert np.max(np.abs(res - rref) / rref) < 1e-12


@pytest.mark.parametrize("b", [20, 150])
def test_project_rotate_residuals(rcg, b):
    """b = 150 spans three 64-column tiles: the projection multiplies only the
    tiles on and below the diagonal and mirrors the lower triangle."""
    print(f"case: projection, rotation and residuals on 12^3 Anderson, b={b}", flush=True)
    A = anderson3d(12)
    S = A.to_scipy()
    op = rcg.CsrOperator(A)
    rng = np.random.default_rng(4)
    V, _ = np.linalg.qr(rng.standard_normal((A.n, b)))
    bv, bw, bt = rcg.Block.from_host(V), rcg.Block(A.n, b), rcg.Block(A.n, b)
    H = np.zeros((b, b), order="F")
    capi.check(capi.lib().rcg_project(op.h, bv.h, bw.h, b, H.ctypes.data_as(C.c_void_p)))
    Href = V.T @ (S @ V)
    assert np.max(np.abs(H - Href)) < 1e-13 * np.max(np.abs(Href))
    assert np.array_equal(H, H.T)
    w, Z = np.linalg.eigh(H)
    Z = np.asfortranarray(Z)
    capi.check(capi.lib().rcg_rotate(bv.h, bt.h, b, Z.ctypes.data_as(C.c_void_p)))
    Vr = bv.download()
    ref = V @ Z
    ref /= np.linalg.norm(ref, axis=0)
    assert np.max(np.abs(Vr - ref)) < 1e-13
    res = np.zeros(b)
    capi.check(capi.lib().rcg_residuals(op.h, bv.h, bw.h, b, w.ctypes.data_as(C.c_void_p),
                                        res.ctypes.data_as(C.c_void_p)))
    rref = np.linalg.norm(w * Vr - S @ Vr, axis=0)
    assert np.max(np.abs(res - rref)) < 1e-12


def test_lanczos_bounds_vs_oracle(rcg, orc):
    print("case: Lanczos bounds 64x64 Dirichlet and 16^3 Anderson", flush=True)
    for A in (ops.laplacian_2d(64), anderson3d(16)):
        op = rcg.CsrOperator(A)
        for flags, (mx_f, mn_f) in ((1, (True, False)), (3, (True, True))):
            cnt = C.c_uint64(0)
            mn, mx, it = C.c_double(), C.c_double(), C.c_int64()
            capi.check(capi.lib().rcg_lanczos_bounds(op.h, 1e-6, flags, 10000, 0, C.byref(cnt),
                                                     C.byref(mn), C.byref(mx), C.byref(it)))
            omn, omx, oit = orc.lanczos_bounds(A, seed=0, tol=1e-6, max_flag=mx_f, min_flag=mn_f)
            assert it.value == oit
            assert cnt.value == A.n
            assert abs(mx.value - omx) < 1e-11 * abs(omx)
            assert abs(mn.value - omn) < 1e-9 * max(1.0, abs(omn))
        # an iteration cap that is not a multiple of the test period (5): the
        # last partial chunk runs untested, as in the reference loop
        for itmax in (7, 12):
            cnt = C.c_uint64(0)
            mn, mx, it = C.c_double(), C.c_double(), C.c_int64()
            capi.check(capi.lib().rcg_lanczos_bounds(op.h, 1e-6, 3, itmax, 0, C.byref(cnt),
                                                     C.byref(mn), C.byref(mx), C.byref(it)))
            omn, omx, oit = orc.lanczos_bounds(A, seed=0, tol=1e-6, max_flag=True, min_flag=True,
                                               itmax=itmax)
            assert it.value == oit == itmax
            assert abs(mx.value - omx) <= 1e-11 * abs(omx)
            assert abs(mn.value - omn) <= 1e-9 * max(1.0, abs(omn))
