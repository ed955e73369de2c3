"""The CPU oracle against what the mathematics fixes.

Closed forms of the scaled Chebyshev polynomial (LanczosCpoly.h:203-232,
363-369), the filter's eigenvector identity, numpy on small dense problems,
analytic Laplacian spectra and the shared random stream.
"""
import numpy as np
import pytest

from rayleighchebyshevgpgpu_amd import operators as ops


def closed_pm(m, x):
    th = 2.0 * np.arcsin(np.sqrt(x))
    return np.sin((m + 1) * th) / ((m + 1) * np.sin(th))


@pytest.mark.parametrize("m", [2, 3, 7, 20, 101, 200])
def test_pm_matches_closed_form(orc, m):
    for x in np.linspace(0.001, 0.999, 41):
        assert abs(orc.cpoly_pm(m, x) - closed_pm(m, x)) < 1e-11


def test_pm_low_degrees(orc):
    for x in np.linspace(-0.5, 1.5, 17):
        assert orc.cpoly_pm(0, x) == 1.0
        assert orc.cpoly_pm(1, x) == pytest.approx(1 - 2 * x, abs=1e-15)
        assert orc.cpoly_pm(2, x) == pytest.approx(1 - 16 * x / 3 + 16 * x * x / 3, abs=1e-13)


@pytest.mark.parametrize("m", [3, 10, 64, 200])
def test_upper_xstar_is_next_to_last_root(orc, m):
    r = orc.cpoly_upper_xstar(m)
    assert r == pytest.approx((np.cos(2 * np.pi / (m + 1)) + 1) / 2, abs=1e-15)
    assert abs(orc.cpoly_pm(m, r)) < 1e-10


@pytest.mark.parametrize("m", [3, 10, 64, 200])
def test_xstar_level_matches_first_interior_extremum(orc, m):
    """|P_m(X*)| equals |P_m| at the first root of P_m' (the secondary maximum)."""
    xp = orc.cpoly_pm_prime_first_root(m)
    assert abs(orc.cpoly_pm_prime(m, xp)) < 1e-6
    xs = orc.cpoly_xstar(m)
    assert 0 < xs < xp
    assert orc.cpoly_pm(m, xs) == pytest.approx(-orc.cpoly_pm(m, xp), rel=1e-8)


@pytest.mark.parametrize("m", [4, 30, 150])
def test_inverse_pm_round_trip(orc, m):
    for y in [0.3, 0.9, 2.0, 4.9, 50.0, 1e6]:
        x = orc.cpoly_inverse_pm(m, y)
        assert orc.cpoly_pm(m, x) == pytest.approx(y, rel=1e-9)


def test_star_degree_maps_lambda_star_to_xstar(orc):
    for shift, lb, ls in [(-0.1, 8.0, 0.5), (0.0, 14.0, 0.01), (2.0, 10.0, -1.5)]:
        m, bound = orc.cpoly_star_degree(shift, lb, ls, 200)
        if m < 200:
            xs, ux = orc.cpoly_xstar(m), orc.cpoly_upper_xstar(m)
            rho = (bound + shift) / ux
            assert (ls + shift) / rho == pytest.approx(xs, rel=1e-12)
            if m > 1:  # m-1 would not reach
                xs1, ux1 = orc.cpoly_xstar(m - 1), orc.cpoly_upper_xstar(m - 1)
                assert (xs1 / ux1) * (lb + shift) >= ls + shift


def test_filter_eigenvector_identity(orc):
    """P_m(A') v = P_m(x(lambda)) v for a 1-D Dirichlet Laplacian eigenvector."""
    n = 200
    A = ops.laplacian_2d(n, 1)  # 1-D chain: diag 4 minus ... use explicit 1-D
    rows = np.arange(n)
    A = ops._from_coo(n, np.concatenate([rows, rows[1:], rows[:-1]]),
                      np.concatenate([rows, rows[:-1], rows[1:]]),
                      np.concatenate([2 * np.ones(n), -np.ones(n - 1), -np.ones(n - 1)]))
    j = np.arange(1, 6)
    i = np.arange(1, n + 1)
    V = np.sin(np.pi * np.outer(i, j) / (n + 1))
    lam = 4 * np.sin(np.pi * j / (2 * (n + 1))) ** 2
    for m, lmax, shift in [(1, 4.0, -lam[0]), (7, 4.01, -lam[0] + 0.001), (40, 4.2, 0.0)]:
        out = orc.filter_csr(A, V, m, lmax, shift)
        for c in range(5):
            p = orc.cpoly_eval_polynomial(m, 1, lmax, shift, lam[c])
            assert np.max(np.abs(out[:, c] - p * V[:, c])) <= 1e-10 * max(1.0, abs(p)) * np.max(np.abs(V[:, c]))


def test_filter_dense_brute_force(orc):
    rng = np.random.default_rng(1)
    n = 60
    G = rng.standard_normal((n, n))
    H = (G + G.T) / 2
    w, U = np.linalg.eigh(H)
    X = rng.standard_normal((n, 4))
    m, lmax, shift = 9, w[-1] + 0.1, -w[0]
    out = orc.filter_dense(H, X, m, lmax, shift)
    p = np.array([orc.cpoly_eval_polynomial(m, 1, lmax, shift, x) for x in w])
    ref = U @ (p[:, None] * (U.T @ X))
    assert np.max(np.abs(out - ref)) < 1e-10 * np.max(np.abs(ref))
    out2 = orc.filter_dense(H, X, m, lmax, shift, rep=2)
    ref2 = U @ ((p ** 2)[:, None] * (U.T @ X))
    assert np.max(np.abs(out2 - ref2)) < 1e-10 * np.max(np.abs(ref2))


def test_tridiagonal_bisection(orc):
    rng = np.random.default_rng(3)
    for n in [1, 2, 5, 40, 200]:
        D = rng.standard_normal(n)
        E = rng.standard_normal(max(n - 1, 0))
        T = np.diag(D) + np.diag(E, 1) + np.diag(E, -1)
        w = np.linalg.eigvalsh(T)
        for idx in {0, n - 1, n // 2}:
            assert orc.tri_eig(D, E, idx) == pytest.approx(w[idx], abs=2e-13)


def test_jacobi_real(orc):
    rng = np.random.default_rng(4)
    for n in [1, 2, 7, 64]:
        G = rng.standard_normal((n, n))
        H = (G + G.T) / 2
        w, V = orc.jacobi(H)
        assert np.allclose(w, np.linalg.eigvalsh(H), atol=1e-12)
        assert np.max(np.abs(H @ V - V * w)) < 1e-12 * max(1, np.max(np.abs(w)))
        assert np.max(np.abs(V.T @ V - np.eye(n))) < 1e-13


def test_jacobi_hermitian(orc):
    rng = np.random.default_rng(5)
    for n in [1, 3, 16, 48]:
        G = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        H = (G + G.conj().T) / 2
        w, V = orc.jacobi(H)
        assert np.allclose(w, np.linalg.eigvalsh(H), atol=1e-12)
        assert np.max(np.abs(H @ V - V * w)) < 1e-12 * max(1, np.max(np.abs(w)))
        assert np.max(np.abs(V.conj().T @ V - np.eye(n))) < 1e-13


def test_mgs_orthonormal_and_triangular(orc):
    rng = np.random.default_rng(6)
    X = rng.standard_normal((500, 12))
    Q = orc.mgs(X)
    assert np.max(np.abs(Q.T @ Q - np.eye(12))) < 1e-13
    R = Q.T @ X
    assert np.max(np.abs(np.tril(R, -1))) < 1e-12 * np.max(np.abs(R))


def test_random_stream_matches_numpy(orc):
    for seed, start in [(0, 0), (1, 12345), (4, 2 ** 40)]:
        a = orc.sym_draws(seed, start, 1000)
        b = ops.sym_draws(seed, start, 1000)
        assert np.array_equal(a, b)
        assert a.min() >= -1 and a.max() < 1
    u = ops.unit_draws(0, 0, 200000)
    assert abs(u.mean() - 0.5) < 5e-3 and abs(u.var() - 1 / 12) < 2e-3


def test_lanczos_bounds_analytic(orc):
    A = ops.laplacian_2d(64)
    exact = ops.laplacian_2d_spectrum(64)
    mn, mx, it = orc.lanczos_bounds(A, seed=0, tol=1e-6, max_flag=True, min_flag=True)
    assert mx == pytest.approx(exact[-1], rel=1e-6)
    assert mn == pytest.approx(exact[0], rel=1e-5)
    assert it % 5 == 0


def test_solver_config1_analytic(orc):
    """64x64 Dirichlet, lowest 16: 4 sin^2(pi i/130) + 4 sin^2(pi j/130)."""
    A = ops.laplacian_2d(64)
    r = orc.solve(A, 16, 1e-10, 8, 8, seed=0, stop="RESIDUAL_ONLY")
    exact = ops.laplacian_2d_spectrum(64)[:16]
    assert r["found"] == 16
    assert np.max(np.abs(r["eigenvalues"] - exact) / np.maximum(exact, 0.1)) < 1e-10
    assert exact[15] == pytest.approx(0.060451911483, abs=1e-11)


def test_solver_periodic3d_analytic(orc):
    """W=0 periodic 3-D Laplacian: 4[sin^2(pi a/n)+sin^2(pi b/n)+sin^2(pi c/n)]."""
    A = ops.laplacian_3d(12, bc="periodic")
    exact = ops.laplacian_3d_spectrum(12)
    r = orc.solve(A, 7, 1e-10, 8, 8, seed=1, stop="RESIDUAL_ONLY", want_vectors=True)
    assert np.max(np.abs(r["eigenvalues"] - exact[:7])) < 1e-9
    V = r["eigenvectors"]
    S = A.to_scipy()
    res = np.linalg.norm(S @ V - V * r["eigenvalues"], axis=0)
    assert np.max(res) < 1e-8


def test_solver_complex_hermitian_small(orc):
    """Hofstadter 16x16, phi=1/8, W=1: eigenvalues against numpy eigvalsh."""
    A = ops.hofstadter_2d(16, phi=1 / 8, W=1.0, seed=2)
    H = A.to_scipy().toarray()
    assert np.allclose(H, H.conj().T)
    exact = np.linalg.eigvalsh(H)
    r = orc.solve(A, 12, 1e-10, 6, 6, seed=0, stop="RESIDUAL_ONLY")
    assert np.max(np.abs(r["eigenvalues"] - exact[:12])) < 1e-9


def test_solver_dense(orc):
    n = 300
    lam = np.linspace(0, 1, n) ** 2
    A = ops.dense_qlq(n, lam, seed=3)
    r = orc.solve(A, 10, 1e-10, 10, 10, seed=0, stop="RESIDUAL_ONLY")
    assert np.max(np.abs(r["eigenvalues"] - lam[:10])) < 1e-10


def test_nonhermitian_operator_raises(orc):
    n = 30
    rows = np.arange(n)
    A = ops._from_coo(n, np.concatenate([rows, rows[:-1]]), np.concatenate([rows, rows[1:]]),
                      np.concatenate([np.ones(n), 1j * np.ones(n - 1)]).astype(complex))
    with pytest.raises(ValueError):
        orc.solve(A, 4, 1e-8, 4, 4)
