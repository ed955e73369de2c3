"""Dense symmetric operators: MFMA block products and whole solves.

Pins: numpy products (FP64 to 1e-13 relative); the FP64 oracle filter on the
same matrix; the known spectrum of A = Q diag(lambda) Q^T with
lambda = linspace(0, 1, n)^2 (the dense workload's construction); the
eigenvectors' residuals recomputed on the host; and, for the FP32 MFMA path,
the mixed filter on exact eigenvectors (the columns of Q), which must come out
at FP64 accuracy because the FP32 steps only carry each column's correction.
"""
import numpy as np
import pytest

from rayleighchebyshevgpgpu_amd import capi
from rayleighchebyshevgpgpu_amd import operators as ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rcg():
    from rayleighchebyshevgpgpu_amd import solver
    return solver


def qlq(n, seed=3):
    lam = np.linspace(0.0, 1.0, n) ** 2
    G = np.random.default_rng(seed).standard_normal((n, n))
    Q, R = np.linalg.qr(G)
    Q = Q * np.sign(np.diag(R))
    A = (Q * lam) @ Q.T
    return np.asfortranarray(0.5 * (A + A.T)), lam, Q


@pytest.mark.parametrize("n,cols", [(300, 13), (1000, 64), (517, 100)])
def test_dense_apply_vs_numpy(rcg, n, cols):
    print(f"case: dense apply n={n} cols={cols}", flush=True)
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    A = np.asfortranarray(G + G.T)
    X = rng.standard_normal((n, cols))
    Z = rng.standard_normal((n, cols))
    op = rcg.DenseOperator(A)
    bx, bz = rcg.Block.from_host(X), rcg.Block.from_host(Z)
    op.apply(bx, bz, 0.5, 2.0)
    ref = 0.5 * (A @ X) + 2.0 * Z
    assert np.max(np.abs(bz.download() - ref)) < 1e-13 * np.max(np.abs(ref)) * np.sqrt(n)


def test_dense_filter_vs_oracle(rcg, orc):
    print("case: dense Chebyshev filter n=600 degree 7, 40 columns", flush=True)
    A, lam, Q = qlq(600)
    X = np.random.default_rng(1).uniform(-1, 1, (600, 40))
    ref = orc.filter_dense(A, X, 7, 1.001, -lam[0])
    op = rcg.DenseOperator(A)
    bx, bt = rcg.Block.from_host(X), rcg.Block(600, 40)
    capi.check(capi.lib().rcg_cheb_filter(op.h, bx.h, bt.h, 40, 7, 1, 1.001, -lam[0]))
    assert np.max(np.abs(bx.download() - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_dense_mixed_filter_exact_eigenvectors(rcg, orc):
    print("case: dense mixed (FP32 MFMA) filter on exact eigenvectors, degree 9", flush=True)
    A, lam, Q = qlq(700)
    X = np.ascontiguousarray(Q[:, :24])
    th = lam[:24].copy()
    shift = -lam[0]
    ref = orc.filter_dense(A, X, 9, 1.001, shift)
    op = rcg.DenseOperator(A)
    bx, bt = rcg.Block.from_host(X), rcg.Block(700, 24)
    ritz = np.ones(24, dtype=np.int32)
    capi.check(capi.lib().rcg_cheb_filter_mixed(op.h, bx.h, bt.h, 24, 9, 1, 1.001, shift,
                                                th.ctypes.data, ritz.ctypes.data))
    got = bx.download()
    scale = np.max(np.abs(ref))
    # exact eigenvectors: the correction is FP64 round-off, so FP32 steps cost nothing
    assert np.max(np.abs(got - ref)) <= 1e-11 * scale
    # plain FP32 columns: FP32 accuracy (the FP32 MFMA path really ran)
    bx2 = rcg.Block.from_host(X)
    capi.check(capi.lib().rcg_cheb_filter_mixed(op.h, bx2.h, bt.h, 24, 9, 1, 1.001, shift,
                                                th.ctypes.data, None))
    err_plain = np.max(np.abs(bx2.download() - ref)) / scale
    print(f"plain FP32 relative error {err_plain:.2e}")
    assert 1e-10 < err_plain < 1e-4


@pytest.mark.parametrize("mixed", [False, True])
def test_dense_solve_known_spectrum(rcg, orc, mixed):
    n, k = 1200, 32
    print(f"case: dense solve n={n} k={k} lambda=linspace(0,1)^2 mixed={mixed}", flush=True)
    A, lam, Q = qlq(n)
    rc = rcg.RayleighChebyshevLM()
    rc.setStopCondition("RESIDUAL_ONLY")
    if mixed:
        rc.filterPrecision = np.float32
    op = rcg.DenseOperator(A)
    V = rcg.Block(n, 48)
    ev = []
    tol = 1e-10
    f = rc.getMinEigenSystem(k, tol, 16, 16, None, op, rcg.RandomizeOp(0), ev, V)
    ev = np.array(ev)
    assert f == k
    floor = np.maximum(np.abs(lam[:k]), 1e-11 / tol)
    err = np.max(np.abs(ev - lam[:k]) / floor)
    print(f"max relative eigenvalue error vs the known spectrum: {err:.2e}")
    assert err <= 1e-10
    X = V.download(f)
    r = np.linalg.norm(A @ X - X * ev, axis=0)  # ||A|| = 1
    print(f"max ||Av - lambda v||: {np.max(r):.2e}")
    assert np.max(r) <= tol
    o = orc.solve(A, k, tol, 16, 16, seed=0, stop="RESIDUAL_ONLY")
    assert np.max(np.abs(ev - o["eigenvalues"]) / floor) <= 1e-10
