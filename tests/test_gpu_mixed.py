"""Mixed-precision filter (FP32 Chebyshev steps on the correction of each
Ritz column, FP64 elsewhere) and mixed-precision solves.

The filter is pinned against the FP64 oracle filter: for a column x with
Ritz value theta and residual r = A x - theta x, the FP32 steps only carry
the correction X_j - u_j(theta) x, so the error of the output is about
eps32 * |r| * (filter growth), far below the eps32 * |x| of a direct FP32
filter.  Exact eigenvectors of the Dirichlet Laplacian (r ~ 1e-16) must come
out at FP64 accuracy; perturbed ones (|r| ~ 1e-4) within 1e-9; plain FP32
columns within 1e-5.  Solves must meet the same bar as FP64 solves:
eigenvalues within 1e-10 of the FP64 oracle (floor max(|lambda|, 1e-11/tol))
and every pair's residual below the requested tolerance.
"""
import numpy as np
import pytest

from rayleighchebyshevgpgpu_amd import capi
from rayleighchebyshevgpgpu_amd import operators as ops
from tests.golden.cases import CASES, build_operator

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rcg():
    from rayleighchebyshevgpgpu_amd import solver
    return solver


def dirichlet_modes(nx, pairs):
    """Normalized eigenvectors and eigenvalues of the nx x nx 5-point Dirichlet Laplacian."""
    t = np.arange(1, nx + 1)
    cols, lam = [], []
    for a, b in pairs:
        sx = np.sin(np.pi * a * t / (nx + 1))
        sy = np.sin(np.pi * b * t / (nx + 1))
        v = np.outer(sy, sx).ravel()  # row r = x + nx*y
        cols.append(v / np.linalg.norm(v))
        lam.append(4 - 2 * np.cos(np.pi * a / (nx + 1)) - 2 * np.cos(np.pi * b / (nx + 1)))
    return np.array(cols).T, np.array(lam)


def mixed_filter(rcg, A, X, m, lmax, shift, theta=None, ritz=None):
    op = rcg.CsrOperator(A)
    cols = X.shape[1]
    bx, bt = rcg.Block.from_host(X), rcg.Block(A.n, cols)
    th = None if theta is None else np.ascontiguousarray(theta, dtype=np.float64)
    rz = None if ritz is None else np.ascontiguousarray(ritz, dtype=np.int32)
    capi.check(capi.lib().rcg_cheb_filter_mixed(
        op.h, bx.h, bt.h, cols, m, 1, lmax, shift,
        None if th is None else th.ctypes.data, None if rz is None else rz.ctypes.data))
    return bx.download()


@pytest.mark.parametrize("m", [2, 3, 15])
def test_mixed_filter_exact_eigenvectors(rcg, orc, m):
    print(f"case: mixed filter, exact Dirichlet 48x48 eigenvectors, degree {m}", flush=True)
    nx = 48
    A = ops.laplacian_2d(nx)
    X, lam = dirichlet_modes(nx, [(1, 1), (1, 2), (2, 1), (2, 2), (3, 1), (1, 3), (3, 2), (2, 3)])
    lmax, shift = 8.0 * 1.001, -lam[0]
    ref = orc.filter_csr(A, X, m, lmax, shift)
    got = mixed_filter(rcg, A, X, m, lmax, shift, lam, np.ones(len(lam)))
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-12 * scale
    # and the filter's scalar on each eigenvector (closed form of the polynomial)
    p = np.array([orc.cpoly_eval_polynomial(m, 1, lmax, shift, l) for l in lam])
    assert np.max(np.abs(got - X * p)) <= 1e-12 * scale


def test_mixed_filter_perturbed_and_plain(rcg, orc):
    print("case: mixed filter, perturbed Ritz columns vs plain FP32 columns, degree 20", flush=True)
    nx = 48
    A = ops.laplacian_2d(nx)
    S = A.to_scipy()
    X0, _ = dirichlet_modes(nx, [(1, 1), (1, 2), (2, 1), (2, 2), (3, 1), (1, 3)])
    rng = np.random.default_rng(5)
    X = X0 + 1e-4 * rng.standard_normal(X0.shape) / np.sqrt(A.n)
    X /= np.linalg.norm(X, axis=0)
    theta = np.einsum("ij,ij->j", X, S @ X)  # Rayleigh quotients
    lmax, shift, m = 8.0 * 1.001, -theta[0], 20
    ref = orc.filter_csr(A, X, m, lmax, shift)
    ritz = mixed_filter(rcg, A, X, m, lmax, shift, theta, np.ones(len(theta)))
    plain = mixed_filter(rcg, A, X, m, lmax, shift, theta, np.zeros(len(theta)))
    scale = np.max(np.abs(ref))
    err_ritz = np.max(np.abs(ritz - ref)) / scale
    err_plain = np.max(np.abs(plain - ref)) / scale
    print(f"relative error: Ritz columns {err_ritz:.2e}, plain FP32 columns {err_plain:.2e}")
    assert err_ritz <= 1e-9
    assert 1e-9 < err_plain <= 1e-5  # plain columns really ran in FP32
    assert err_ritz * 100 < err_plain


def test_mixed_filter_degree1_is_fp64(rcg, orc):
    print("case: mixed filter degree 1 (FP64 step)", flush=True)
    A = build_operator(CASES["anderson3d16_k32"])
    X = np.random.default_rng(2).uniform(-1, 1, (A.n, 10))
    ref = orc.filter_csr(A, X, 1, 14.5, 1.7)
    got = mixed_filter(rcg, A, X, 1, 14.5, 1.7)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def mixed_solve(rcg, A, c, want_vectors=False):
    rc = rcg.RayleighChebyshevLM()
    rc.setStopCondition(c.get("stop", "RESIDUAL_ONLY"))
    rc.filterPrecision = np.float32
    op = rcg.CsrOperator(A)
    ev = []
    V = rcg.Block(A.n, max(c["k"], c["inc"] + c["buf"])) if want_vectors else None
    f = rc.getMinEigenSystem(c["k"], c["tol"], c["inc"], c["buf"], None, op,
                             rcg.RandomizeOp(c.get("seed", 0)), ev, V)
    return rc, f, np.array(ev), V


def test_mixed_solve_config1_analytic(rcg):
    print("case: mixed solve, 64x64 Dirichlet k=16 inc=8 buf=8 tol 1e-10", flush=True)
    A = ops.laplacian_2d(64)
    c = dict(k=16, inc=8, buf=8, tol=1e-10, seed=0)
    rc, f, ev, V = mixed_solve(rcg, A, c, want_vectors=True)
    exact = ops.laplacian_2d_spectrum(64)[:16]
    assert f == 16
    assert np.max(np.abs(ev - exact) / np.maximum(exact, 0.1)) < 1e-10
    X = V.download(16)
    r = np.linalg.norm(A.to_scipy() @ X - X * ev, axis=0)
    assert np.max(r) / 8.0 < 1e-10
    assert np.max(np.abs(X.T @ X - np.eye(16))) < 1e-10


@pytest.mark.parametrize("name", ["anderson3d16_k32", "lap2d40_k12_eig", "anderson2d48_k16"])
def test_mixed_solve_matches_fp64_oracle(rcg, orc, name):
    c = dict(CASES[name])
    c["stop"] = "RESIDUAL_ONLY"
    c["tol"] = 1e-10
    print(f"case: mixed solve {name} {c}", flush=True)
    A = build_operator(CASES[name])
    rc, f, ev, V = mixed_solve(rcg, A, c, want_vectors=True)
    o = orc.solve(A, c["k"], c["tol"], c["inc"], c["buf"], seed=c.get("seed", 0),
                  stop="RESIDUAL_ONLY")
    assert f == o["found"]
    floor = np.maximum(np.abs(o["eigenvalues"]), 1e-11 / c["tol"])
    err = np.abs(ev - o["eigenvalues"]) / floor
    print(f"max relative eigenvalue error vs FP64 oracle: {np.max(err):.2e}")
    assert np.all(err <= 1e-10)
    X = V.download(f)
    S = A.to_scipy()
    normA = max(abs(rc.getMaxEigValueEst()), abs(rc.getMinEigValueEst()))  # ~ ||A||_2
    r = np.linalg.norm(S @ X - X * ev, axis=0) / normA
    print(f"max ||Av - lambda v||/||A||: {np.max(r):.2e}")
    assert np.max(r) <= c["tol"]
