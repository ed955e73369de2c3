"""Caller-supplied operators (rcg_operator_callback_create): the reference's
Otype::apply(in, out, alpha, beta) contract (LanczosCpolyOperatorLMCuda.h
:236,245, RayleighChebyshevLMCuda.h:2135,2235) as compiled matrix-free code.

The operators (tests/cpp/stencil_ops.hip) never build a matrix: the 3-D
periodic Laplacian plus an Anderson potential and the 2-D Hofstadter lattice
are stencil kernels on column-major blocks.  Pins: the product equals the CSR
operator's (1e-13 of its largest entry); the Chebyshev filter through the
caller's apply equals the oracle filter (LanczosCpolyOperatorLM.h:192-304) on
the same operator's CSR to 1e-12; whole FP64 solves (real and complex) agree
with the oracle's eigenvalues to 1e-10 with the reference's floor and find the
same counts, also with the apply called through a Python function; mixed
precision and a failing apply are reported as errors."""
import ctypes as C
import os

import numpy as np
import pytest

from rayleighchebyshevgpgpu_amd import capi
from rayleighchebyshevgpgpu_amd import operators as ops

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def rcg():
    from rayleighchebyshevgpgpu_amd import solver
    return solver


@pytest.fixture(scope="module")
def stencils():
    import __graft_entry__ as ge
    src = os.path.join(ROOT, "tests", "cpp", "stencil_ops.hip")
    if not os.path.exists(ge.STENCIL_LIB) or os.path.getmtime(ge.STENCIL_LIB) < os.path.getmtime(src):
        ge.build_stencil_lib()
    capi.lib()  # librcg first: both use the system HIP runtime
    L = C.CDLL(ge.STENCIL_LIB)
    L.anderson3d_create.restype = C.c_void_p
    L.anderson3d_create.argtypes = [C.c_int64, C.c_void_p]
    L.hofstadter_create.restype = C.c_void_p
    L.hofstadter_create.argtypes = [C.c_int64, C.c_double, C.c_void_p]
    for f in (L.anderson3d_destroy, L.hofstadter_destroy):
        f.argtypes = [C.c_void_p]
    for f in (L.anderson3d_apply, L.hofstadter_apply):
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64,
                      C.c_double, C.c_double, C.c_void_p]
        f.restype = C.c_int
    return L


def anderson(rcg, stencils, L, seed=1, via_python=False):
    pot = ops.anderson_potential(L ** 3, 4.0, seed)
    A = ops.laplacian_3d(L, bc="periodic", potential=pot)
    diag = np.ascontiguousarray(6.0 + pot)
    user = stencils.anderson3d_create(L, diag.ctypes.data)
    assert user
    fn = C.cast(stencils.anderson3d_apply, C.c_void_p).value
    if via_python:
        def apply(pin, pout, n, cols, ld, alpha, beta, stream):
            rc = stencils.anderson3d_apply(user, pin, pout, n, cols, ld, alpha, beta, stream)
            if rc != 0:
                raise RuntimeError(f"stencil apply failed ({rc})")
        op = rcg.CallbackOperator(A.n, apply)
    else:
        op = rcg.CallbackOperator.from_c(A.n, fn, user)
    return A, op, user


def hofstadter(rcg, stencils, L, phi):
    H = ops.hofstadter_2d(L, phi=phi, W=1.0, seed=2)
    onsite = np.ascontiguousarray(H.to_scipy().diagonal().real)
    user = stencils.hofstadter_create(L, phi, onsite.ctypes.data)
    assert user
    fn = C.cast(stencils.hofstadter_apply, C.c_void_p).value
    return H, rcg.CallbackOperator.from_c(H.n, fn, user, dtype=np.complex128), user


def test_callback_product_matches_csr(rcg, stencils):
    print("case: caller apply vs CSR apply, 10^3 Anderson and 32^2 Hofstadter", flush=True)
    A, aop, u1 = anderson(rcg, stencils, 10)
    H, hop, u2 = hofstadter(rcg, stencils, 32, 1.0 / 8)
    rng = np.random.default_rng(3)
    for M, cop, dt in ((A, aop, np.float64), (H, hop, np.complex128)):
        cols = 7
        X = rng.standard_normal((M.n, cols))
        Y0 = rng.standard_normal((M.n, cols))
        if dt == np.complex128:
            X = X + 1j * rng.standard_normal((M.n, cols))
            Y0 = Y0 + 1j * rng.standard_normal((M.n, cols))
        bx, by1, by2 = rcg.Block.from_host(X), rcg.Block.from_host(Y0), rcg.Block.from_host(Y0)
        rcg.CsrOperator(M).apply(bx, by1, 0.7, -0.3, cols)
        cop.apply(bx, by2, 0.7, -0.3, cols)
        ref = 0.7 * (M.to_scipy() @ X) - 0.3 * Y0
        got1, got2 = by1.download(), by2.download()
        assert np.max(np.abs(got2 - ref)) <= 1e-13 * np.max(np.abs(ref))
        assert np.max(np.abs(got2 - got1)) <= 1e-13 * np.max(np.abs(ref))
    del aop, hop
    stencils.anderson3d_destroy(u1)
    stencils.hofstadter_destroy(u2)


@pytest.mark.parametrize("m,cols", [(1, 3), (6, 16), (25, 40)])
def test_callback_filter_vs_oracle(rcg, orc, stencils, m, cols):
    print(f"case: filter through the caller's apply, degree {m}, {cols} columns", flush=True)
    A, cop, user = anderson(rcg, stencils, 12)
    X = np.random.default_rng(m).uniform(-1, 1, (A.n, cols))
    lmax, shift = 14.5, 1.7
    ref = orc.filter_csr(A, X, m, lmax, shift)
    bx, bt = rcg.Block.from_host(X), rcg.Block(A.n, cols)
    capi.check(capi.lib().rcg_cheb_filter(cop.h, bx.h, bt.h, cols, m, 1, lmax, shift))
    got = bx.download()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    del cop
    stencils.anderson3d_destroy(user)


@pytest.mark.parametrize("via_python", [False, True])
def test_callback_solve_real_vs_oracle(rcg, orc, stencils, via_python):
    print(f"case: solve through a matrix-free 12^3 Anderson apply, k 16 (python={via_python})",
          flush=True)
    k, inc, buf, tol = 16, 8, 8, 1e-10
    A, cop, user = anderson(rcg, stencils, 12, via_python=via_python)
    rc = rcg.RayleighChebyshevLM()
    rc.setStopCondition("RESIDUAL_ONLY")
    ev = []
    V = rcg.Block(A.n, k)
    found = rc.getMinEigenSystem(k, tol, inc, buf, None, cop, rcg.RandomizeOp(0), ev, V)
    o = orc.solve(A, k, tol, inc, buf, seed=0, stop="RESIDUAL_ONLY")
    ev = np.array(ev)
    assert found == o["found"] == k
    assert rc.getResult().lanczos_iters == o["lanczos_iters"]
    floor = np.maximum(np.abs(o["eigenvalues"]), 1e-11 / tol)
    assert np.max(np.abs(ev - o["eigenvalues"]) / floor) < 1e-10
    X = V.download(k)
    r = np.linalg.norm(A.to_scipy() @ X - X * ev, axis=0)
    assert np.max(r) / 14.0 < tol
    del cop
    stencils.anderson3d_destroy(user)


def test_callback_solve_complex_vs_oracle(rcg, orc, stencils):
    print("case: solve through a matrix-free 32^2 Hofstadter apply, k 12", flush=True)
    k, tol = 12, 1e-10
    H, cop, user = hofstadter(rcg, stencils, 32, 1.0 / 8)
    rc = rcg.RayleighChebyshevLM()
    rc.setStopCondition("RESIDUAL_ONLY")
    ev = []
    V = rcg.Block(H.n, k, dtype=np.complex128)
    found = rc.getMinEigenSystem(k, tol, k, 6, None, cop, rcg.RandomizeOp(0), ev, V)
    o = orc.solve(H, k, tol, k, 6, seed=0, stop="RESIDUAL_ONLY")
    ev = np.array(ev)
    assert found == o["found"] == k
    floor = np.maximum(np.abs(o["eigenvalues"]), 1e-11 / tol)
    assert np.max(np.abs(ev - o["eigenvalues"]) / floor) < 1e-10
    X = V.download(k)
    r = np.linalg.norm(H.to_scipy() @ X - X * ev, axis=0)
    assert np.max(r) / 8.0 < tol
    del cop
    stencils.hofstadter_destroy(user)


def test_callback_errors(rcg, stencils):
    print("case: caller operator errors: mixed precision, a failing apply", flush=True)
    A, cop, user = anderson(rcg, stencils, 8)
    rc = rcg.RayleighChebyshevLM()
    rc.filterPrecision = np.float32
    with pytest.raises(capi.RcgError) as e:
        rc.getMinEigenSystem(4, 1e-8, 4, 4, None, cop, rcg.RandomizeOp(0), [], None)
    assert e.value.code == 6

    def failing(pin, pout, n, cols, ld, alpha, beta, stream):
        raise ValueError("operator refused")
    bad = rcg.CallbackOperator(A.n, failing)
    rc = rcg.RayleighChebyshevLM()
    with pytest.raises(capi.RcgError) as e:
        rc.getMinEigenSystem(4, 1e-8, 4, 4, None, bad, rcg.RandomizeOp(0), [], None)
    assert e.value.code == 1 and isinstance(e.value.__cause__, ValueError)
    assert "callback" in str(e.value)
    del cop
    stencils.anderson3d_destroy(user)
