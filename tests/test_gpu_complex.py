"""Complex Hermitian operators (C128 CSR): device kernels and whole solves.

The operator is the 2-D Hofstadter lattice of the complex-Hermitian workload
(periodic L x L, hopping -1 on x-bonds, -exp(2 pi i phi x) on y-bonds,
onsite U[-W/2, W/2]).  Pins: the CPU oracle's complex solve (eigenvalues to
a relative 1e-10 with the floor max(|lambda|, 1e-11/tol)); numpy eigvalsh of
the dense matrix; residuals recomputed on the host from the downloaded
eigenvectors; the start-vector draws bit for bit (two draws per complex
entry, re then im, ArandomizeOpClassLMCuda.h:66-70, 82-86); and the Lanczos
Hermitian check (LanczosMaxMinFinder.h:350-356) on a non-Hermitian matrix.
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


def test_complex_randomize_bit_exact(rcg):
    print("case: complex randomize cols 1..3 of 500x5, seed 4, counter 10", flush=True)
    n = 500
    b = rcg.Block(n, 5, dtype=np.complex128)
    c = b.randomize(1, 3, 4, 10)
    assert c == 10 + 2 * 3 * n
    X = b.download()
    d = ops.sym_draws(4, 10, 2 * 3 * n)
    for j in range(3):
        re = d[2 * j * n: 2 * (j + 1) * n: 2]
        im = d[2 * j * n + 1: 2 * (j + 1) * n: 2]
        assert np.array_equal(X[:, 1 + j], re + 1j * im)
    assert np.all(X[:, 0] == 0) and np.all(X[:, 4] == 0)


def test_complex_apply_vs_scipy(rcg):
    print("case: complex CSR apply, Hofstadter 40x40 phi=1/8 W=1, 13 columns", flush=True)
    A = ops.hofstadter_2d(40, phi=1 / 8, W=1.0, seed=2)
    op = rcg.CsrOperator(A)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((A.n, 13)) + 1j * rng.standard_normal((A.n, 13))
    Z = rng.standard_normal((A.n, 13)) + 1j * rng.standard_normal((A.n, 13))
    bx, bz = rcg.Block.from_host(X), rcg.Block.from_host(Z)
    op.apply(bx, bz, 0.5, -2.0)
    ref = 0.5 * (A.to_scipy() @ X) - 2.0 * Z
    assert np.max(np.abs(bz.download() - ref)) < 1e-13 * np.max(np.abs(ref))


def complex_solve(rcg, A, k, inc, buf, tol, mixed=False, seed=0, stop="RESIDUAL_ONLY"):
    rc = rcg.RayleighChebyshevLM()
    rc.setStopCondition(stop)
    if mixed:
        rc.filterPrecision = np.float32
    op = rcg.CsrOperator(A)
    V = rcg.Block(A.n, max(k, inc + buf), dtype=np.complex128)
    ev = []
    f = rc.getMinEigenSystem(k, tol, inc, buf, None, op, rcg.RandomizeOp(seed), ev, V)
    return rc, f, np.array(ev), V.download(f)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("L,phi,W,k,inc,buf", [(32, 1 / 8, 1.0, 16, 8, 8), (48, 1 / 16, 2.0, 24, 12, 8),
                                                (48, 1 / 16, 2.0, 40, 24, 16)])
def test_complex_solve_vs_oracle(rcg, orc, L, phi, W, k, inc, buf, mixed):
    print(f"case: complex solve Hofstadter {L}x{L} phi={phi} W={W} k={k} mixed={mixed}", flush=True)
    tol = 1e-10
    A = ops.hofstadter_2d(L, phi=phi, W=W, seed=2)
    rc, f, ev, X = complex_solve(rcg, A, k, inc, buf, tol, mixed=mixed)
    o = orc.solve(A, k, tol, inc, buf, seed=0, stop="RESIDUAL_ONLY")
    assert f == o["found"] == k
    floor = np.maximum(np.abs(o["eigenvalues"]), 1e-11 / tol)
    err = np.max(np.abs(ev - o["eigenvalues"]) / floor)
    print(f"max relative eigenvalue error vs oracle: {err:.2e}")
    assert err <= 1e-10
    S = A.to_scipy()
    normA = max(abs(rc.getMaxEigValueEst()), abs(rc.getMinEigValueEst()))
    r = np.linalg.norm(S @ X - X * ev, axis=0) / normA
    print(f"max ||Av - lambda v||/||A||: {np.max(r):.2e}")
    assert np.max(r) <= tol
    assert np.max(np.abs(X.conj().T @ X - np.eye(f))) < 1e-10
    # the Lanczos bounds come from the same draws as the oracle's
    assert rc.getSpectralBoundsIterCount() == o["lanczos_iters"]


def test_complex_solve_vs_dense_spectrum(rcg):
    print("case: complex solve Hofstadter 32x32 phi=1/8 W=0 vs numpy eigvalsh", flush=True)
    A = ops.hofstadter_2d(32, phi=1 / 8, W=0.0)
    exact = np.linalg.eigvalsh(A.to_scipy().toarray())
    rc, f, ev, X = complex_solve(rcg, A, 10, 8, 8, 1e-10)
    assert f == 10
    assert np.max(np.abs(ev - exact[:10]) / np.maximum(np.abs(exact[:10]), 0.1)) < 1e-10


def test_complex_nonhermitian_raises(rcg):
    print("case: non-Hermitian complex operator -> RCG_ERR_NOT_HERMITIAN", flush=True)
    n = 200
    rows = np.arange(n)
    A = ops._from_coo(n, np.concatenate([rows, rows[:-1]]), np.concatenate([rows, rows[1:]]),
                      np.concatenate([np.ones(n), 1j * np.ones(n - 1)]).astype(complex))
    with pytest.raises(capi.RcgError) as e:
        complex_solve(rcg, A, 4, 4, 4, 1e-8)
    assert e.value.code == 4
