"""The code paths the benchmarks run, on operators small enough for the oracle.

* Every row-block visiting order and every column-panel width the filter
  autotuner can pick, for the four step kinds (FP64, FP32 correction, C128,
  complex64 correction).  The order and the panels change which rows and
  columns are in flight together, never the arithmetic of a column: every
  choice must give bit-identical output, and the default choice matches the
  oracle filter (LanczosCpolyOperatorLM.h:192-304).
* The autotuner itself (it engages when rows x columns >= 2^22).
* Every width-selected instance of the CSR step (16-byte vectors per row
  1..256, i.e. lanes per row and vectors per lane), in FP64, in the FP32
  correction form at the mixed workload's 64 columns and in the complex64 form
  at the complex workload's 96 complex columns.
* Columns past `cols` are never written.
* Complex (C128) blocks through the piecewise entry points.
* Solves of the benchmark shapes: the 32^3 Anderson golden case with the
  128^3 workload's inc 32 / buf 32 / k 64; a complex lattice with the complex
  workload's k 128 / inc 64 / buf 32 (FP64 and mixed) against the oracle; a
  dense solve with the dense workload's inc 128 / buf 128 against the known
  spectrum.

Tolerances: FP64 output within 1e-12 of the oracle (relative to the largest
entry); plain FP32 columns within 1e-5; exact eigenvectors through the
correction form at FP64 accuracy (1e-11); eigenvalues within 1e-10 with the
reference's floor max(|lambda|, 1e-11/tol).
"""
import contextlib
import json
import os

import numpy as np
import pytest

from rayleighchebyshevgpgpu_amd import capi
from rayleighchebyshevgpgpu_amd import operators as ops
from tests.golden.cases import CASES, build_operator

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_cases.json")))

SCHEDULES = ["rows", "rows-noremap"] + [
    f"tile{t}{r}" for t in ("8x8", "16x16", "32x8", "16x4", "4x16") for r in ("", "-noremap")]


@pytest.fixture(scope="module")
def rcg():
    from rayleighchebyshevgpgpu_amd import solver
    return solver


@contextlib.contextmanager
def forced_shape(sched=None, panel=None):
    """Force the filter's row-block order and column panel (the autotuner's
    two choices) through RCG_FILTER_SCHED / RCG_FILTER_PANEL."""
    old = {k: os.environ.get(k) for k in ("RCG_FILTER_SCHED", "RCG_FILTER_PANEL")}
    try:
        if sched is not None:
            os.environ["RCG_FILTER_SCHED"] = sched
        if panel is not None:
            os.environ["RCG_FILTER_PANEL"] = str(panel)
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def anderson3d(nx, W=4.0, seed=1):
    pot = ops.anderson_potential(nx ** 3, W, seed)
    return ops.laplacian_3d(nx, bc="periodic", potential=pot)


def plane_waves(L, modes):
    """Exact eigenvectors of the periodic L x L lattice with hopping -1 (the
    phase-free Hofstadter lattice): exp(2 pi i (a x + b y) / L), eigenvalue
    -2 cos(2 pi a / L) - 2 cos(2 pi b / L); rows numbered x + L y."""
    x = np.tile(np.arange(L), L)
    y = np.repeat(np.arange(L), L)
    cols, lam = [], []
    for a, b in modes:
        cols.append(np.exp(2j * np.pi * (a * x + b * y) / L) / L)
        lam.append(-2 * np.cos(2 * np.pi * a / L) - 2 * np.cos(2 * np.pi * b / L))
    return np.array(cols).T, np.array(lam)


def dirichlet_modes(nx, count):
    t = np.arange(1, nx + 1)
    pairs = sorted(((a, b) for a in range(1, nx + 1) for b in range(1, nx + 1)),
                   key=lambda p: np.sin(np.pi * p[0] / (2 * nx + 2)) ** 2
                   + np.sin(np.pi * p[1] / (2 * nx + 2)) ** 2)[:count]
    cols, lam = [], []
    for a, b in pairs:
        v = np.outer(np.sin(np.pi * b * t / (nx + 1)), np.sin(np.pi * a * t / (nx + 1))).ravel()
        cols.append(v / np.linalg.norm(v))
        lam.append(4 - 2 * np.cos(np.pi * a / (nx + 1)) - 2 * np.cos(np.pi * b / (nx + 1)))
    return np.array(cols).T, np.array(lam)


def run_filter(rcg, op, X, m, lmax, shift, mixed=False, theta=None, ritz=None, cols=None):
    cols = X.shape[1] if cols is None else cols
    bx, bt = rcg.Block.from_host(X), rcg.Block(X.shape[0], X.shape[1], dtype=X.dtype)
    if mixed:
        th = np.zeros(cols) if theta is None else np.ascontiguousarray(theta, dtype=np.float64)
        rz = None if ritz is None else np.ascontiguousarray(ritz, dtype=np.int32)
        capi.check(capi.lib().rcg_cheb_filter_mixed(op.h, bx.h, bt.h, cols, m, 1, lmax, shift,
                                                    th.ctypes.data,
                                                    None if rz is None else rz.ctypes.data))
    else:
        capi.check(capi.lib().rcg_cheb_filter(op.h, bx.h, bt.h, cols, m, 1, lmax, shift))
    return bx.download()


KINDS = {
    # kind: (complex operator, mixed filter, block columns)
    "f64": (False, False, 64),
    "f32src": (False, True, 64),
    "c128": (True, False, 32),
    "c64src": (True, True, 32),
}


@pytest.mark.parametrize("kind", list(KINDS))
def test_filter_every_schedule_and_panel(rcg, orc, kind):
    cx, mixed, cols = KINDS[kind]
    print(f"case: filter kind {kind}, {cols} columns, 12 row orders x 3 panels", flush=True)
    if cx:
        A = ops.hofstadter_2d(64, phi=1 / 16, W=1.0, seed=2)
    else:
        A = anderson3d(16)
    # a locality hint of 64-row x-lines on an 8 x 8 (y, z) grid: 4096 rows
    op = rcg.CsrOperator(A, grid=(64, 8, 8))
    rng = np.random.default_rng(11)
    X = rng.uniform(-1, 1, (A.n, cols))
    if cx:
        X = X + 1j * rng.uniform(-1, 1, (A.n, cols))
    m, lmax, shift = 9, 8.5 if cx else 14.5, 1.7
    with forced_shape("rows", 2 * cols if cx else cols):
        base = run_filter(rcg, op, X, m, lmax, shift, mixed=mixed)
    ref = orc.filter_csr(A, X, m, lmax, shift)
    err = np.max(np.abs(base - ref)) / np.max(np.abs(ref))
    print(f"default order vs oracle: {err:.2e}")
    assert err <= (1e-5 if mixed else 1e-12)
    real_cols = 2 * cols if cx else cols
    for sched in SCHEDULES:
        for panel in (real_cols, 32, 16):
            with forced_shape(sched, panel):
                got = run_filter(rcg, op, X, m, lmax, shift, mixed=mixed)
            assert np.array_equal(got, base), (sched, panel)


def test_unknown_schedule_is_an_error(rcg):
    print("case: RCG_FILTER_SCHED naming no schedule", flush=True)
    A = anderson3d(8)
    op = rcg.CsrOperator(A)  # no grid hint: no tile orders
    X = np.ones((A.n, 4))
    with forced_shape("tile8x8"):
        with pytest.raises(capi.RcgError) as e:
            run_filter(rcg, op, X, 3, 14.0, 1.0)
    assert e.value.code == 1  # RCG_ERR_INVALID


@pytest.mark.parametrize("mixed", [False, True])
def test_autotuned_filter_vs_oracle(rcg, orc, mixed):
    print(f"case: autotuned filter 64^3 Anderson, 16 columns (n*b = 2^22), mixed={mixed}",
          flush=True)
    A = anderson3d(64)
    op = rcg.CsrOperator(A, grid=(64, 64, 64))
    X = np.random.default_rng(3).uniform(-1, 1, (A.n, 16))
    m, lmax, shift = 4, 14.5, 1.9
    ref = orc.filter_csr(A, X, m, lmax, shift)
    got = run_filter(rcg, op, X, m, lmax, shift, mixed=mixed)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"tuned filter vs oracle: {err:.2e}")
    assert err <= (1e-5 if mixed else 1e-12)
    # a second call reuses the cached choice
    again = run_filter(rcg, op, X, m, lmax, shift, mixed=mixed)
    assert np.array_equal(again, got)


FUSED_CASES = {
    # case: (operator, complex, block columns)
    "real5": ("anderson16", False, 5),
    "real13": ("anderson16", False, 13),
    "real64": ("anderson16", False, 64),
    "real100": ("anderson16", False, 100),
    "cx3": ("hofstadter32", True, 3),
    "cx17": ("hofstadter32", True, 17),
    "tuned16": ("anderson64", False, 16),
}


@pytest.mark.parametrize("case", list(FUSED_CASES))
def test_mixed_fused_start_bit_identical(rcg, case):
    """The mixed filter's start computed in the FP64 apply's epilogue (the
    default for CSR operators) against the apply and the start kernel run
    separately (RCG_MIXED_UNFUSED=1): bit-identical outputs, for Ritz and
    plain columns, odd widths (a partly used last vector), complex blocks and
    the autotuned schedule (64^3 x 16 columns)."""
    name, cx, cols = FUSED_CASES[case]
    print(f"case: fused mixed start vs separate kernels, {case}", flush=True)
    if name == "hofstadter32":
        A = ops.hofstadter_2d(32, phi=1 / 8, W=1.0, seed=2)
        lmax, shift = 8.5, 1.7
    else:
        A = anderson3d(16 if name == "anderson16" else 64)
        lmax, shift = 14.5, 1.9
    op = rcg.CsrOperator(A)
    rng = np.random.default_rng(7)
    X = rng.uniform(-1, 1, (A.n, cols))
    if cx:
        X = X + 1j * rng.uniform(-1, 1, (A.n, cols))
    theta = rng.uniform(-1.0, 1.0, cols)
    ritz = (np.arange(cols) % 3 != 2).astype(np.int32)
    fused = run_filter(rcg, op, X, 7, lmax, shift, mixed=True, theta=theta, ritz=ritz)
    old = os.environ.get("RCG_MIXED_UNFUSED")
    os.environ["RCG_MIXED_UNFUSED"] = "1"
    try:
        apart = run_filter(rcg, op, X, 7, lmax, shift, mixed=True, theta=theta, ritz=ritz)
    finally:
        if old is None:
            os.environ.pop("RCG_MIXED_UNFUSED", None)
        else:
            os.environ["RCG_MIXED_UNFUSED"] = old
    assert np.array_equal(fused, apart)


@pytest.mark.parametrize("kind", ["real", "complex", "mixed"])
def test_paired_gram_bit_identical(rcg, kind):
    """The projection's two Gram products (V^H V and V^H AV) in one paired launch
    (the default below 2^22 block entries) against two separate launches
    (RCG_GRAM_UNPAIRED=1): the whole solve, eigenvalues and eigenvectors, is
    bit-identical."""
    print(f"case: paired vs separate Gram products, {kind} solve", flush=True)
    if kind == "complex":
        A = ops.hofstadter_2d(24, phi=1 / 8, W=1.0, seed=2)
        k, inc, buf = 12, 8, 8
    else:
        A = anderson3d(14)
        k, inc, buf = 24, 16, 16
    op = rcg.CsrOperator(A)

    def solve():
        rc = rcg.RayleighChebyshevLM()
        rc.setStopCondition("RESIDUAL_ONLY")
        if kind == "mixed":
            rc.filterPrecision = np.float32
        V = rcg.Block(A.n, inc + buf, dtype=np.complex128 if A.is_complex else np.float64)
        ev = []
        f = rc.getMinEigenSystem(k, 1e-10, inc, buf, None, op, rcg.RandomizeOp(0), ev, V)
        return f, np.array(ev), V.download(f)

    f1, ev1, X1 = solve()
    old = os.environ.get("RCG_GRAM_UNPAIRED")
    os.environ["RCG_GRAM_UNPAIRED"] = "1"
    try:
        f2, ev2, X2 = solve()
    finally:
        if old is None:
            os.environ.pop("RCG_GRAM_UNPAIRED", None)
        else:
            os.environ["RCG_GRAM_UNPAIRED"] = old
    assert f1 == f2 == k
    assert np.array_equal(ev1, ev2)
    assert np.array_equal(X1, X2)


# FP64 widths reaching every (lanes per row, vectors per lane) instance:
# 16-byte vectors per row 1, 2, 3-4, 5-8, 9-16, 17-32, 33-64, 65-128, 129-256
# (and partly used vectors in each)
@pytest.mark.parametrize("cols", [1, 4, 5, 7, 11, 16, 23, 30, 40, 64, 90, 100, 130, 190, 256,
                                  300, 380, 512])
def test_fp64_step_every_width(rcg, orc, cols):
    print(f"case: FP64 filter degree 6, {cols} columns", flush=True)
    A = anderson3d(12)
    op = rcg.CsrOperator(A)
    X = np.random.default_rng(cols).uniform(-1, 1, (A.n, cols))
    ref = orc.filter_csr(A, X, 6, 14.5, 1.7)
    got = run_filter(rcg, op, X, 6, 14.5, 1.7)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("cols", [64, 100, 256])
def test_mixed_step_wide_real(rcg, orc, cols):
    """The FP32 correction step at the 128^3 workload's 64 columns (and wider):
    exact Dirichlet eigenvectors come out at FP64 accuracy, plain FP32 columns
    at FP32 accuracy."""
    print(f"case: mixed filter, {cols} exact Dirichlet 40x40 eigenvectors, degree 12", flush=True)
    nx = 40
    A = ops.laplacian_2d(nx)
    op = rcg.CsrOperator(A)
    X, lam = dirichlet_modes(nx, cols)
    lmax, shift, m = 8.0 * 1.001, -lam[0], 12
    ref = orc.filter_csr(A, X, m, lmax, shift)
    scale = np.max(np.abs(ref))
    got = run_filter(rcg, op, X, m, lmax, shift, mixed=True, theta=lam, ritz=np.ones(cols))
    assert np.max(np.abs(got - ref)) <= 1e-11 * scale
    plain = run_filter(rcg, op, X, m, lmax, shift, mixed=True)
    err = np.max(np.abs(plain - ref)) / scale
    print(f"plain FP32 columns: {err:.2e}")
    assert 1e-10 < err <= 1e-5


@pytest.mark.parametrize("cols", [48, 96])
def test_mixed_step_wide_complex(rcg, orc, cols):
    """The complex64 correction step at the complex workload's 96 complex
    columns (48 and 96: 24 and 48 16-byte vectors per row)."""
    print(f"case: complex mixed filter, {cols} exact plane waves on 64x64, degree 10", flush=True)
    L = 64
    A = ops.hofstadter_2d(L, phi=0.0, W=0.0)
    op = rcg.CsrOperator(A)
    modes = [(a, b) for a in range(-6, 7) for b in range(-6, 7)][:cols]
    X, lam = plane_waves(L, modes)
    lmax, shift, m = 4.0 * 1.001, 4.0, 10
    ref = orc.filter_csr(A, X, m, lmax, shift)
    scale = np.max(np.abs(ref))
    got = run_filter(rcg, op, X, m, lmax, shift, mixed=True, theta=lam, ritz=np.ones(cols))
    assert np.max(np.abs(got - ref)) <= 1e-11 * scale
    # the filter's scalar on each plane wave
    p = np.array([orc.cpoly_eval_polynomial(m, 1, lmax, shift, v) for v in lam])
    assert np.max(np.abs(got - X * p)) <= 1e-11 * scale
    plain = run_filter(rcg, op, X, m, lmax, shift, mixed=True)
    err = np.max(np.abs(plain - ref)) / scale
    print(f"plain complex64 columns: {err:.2e}")
    assert 1e-10 < err <= 1e-5


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_columns_past_cols_untouched(rcg, dtype):
    """Odd column counts end inside a 16-byte vector: the filter steps and the
    apply must leave the block's next column as it was."""
    cx = dtype == np.complex128
    print(f"case: 3 of 4 columns, filter / mixed filter / apply, complex={cx}", flush=True)
    A = ops.hofstadter_2d(24, phi=1 / 8, W=1.0) if cx else anderson3d(10)
    op = rcg.CsrOperator(A)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((A.n, 4)).astype(dtype)
    if cx:
        X = X + 1j * rng.standard_normal((A.n, 4))
    for mixed in (False, True):
        for m in (1, 2, 5):
            got = run_filter(rcg, op, X, m, 9.0, 1.0, mixed=mixed, cols=3)
            assert np.array_equal(got[:, 3], X[:, 3]), (mixed, m)
    bx, by = rcg.Block.from_host(X), rcg.Block.from_host(X[:, ::-1].copy())
    before = by.download()
    op.apply(bx, by, 1.0, 0.0, cols=3)
    assert np.array_equal(by.download()[:, 3], before[:, 3])
    op.apply(bx, by, 2.0, 0.5, cols=3)
    assert np.array_equal(by.download()[:, 3], before[:, 3])


# --------------------------------------------------- complex piecewise
def test_complex_piecewise_entry_points(rcg, orc):
    print("case: C128 filter, CholQR, deflation, projection, rotation on Hofstadter 40x40",
          flush=True)
    L = 40
    A = ops.hofstadter_2d(L, phi=1 / 8, W=1.0, seed=2)
    S = A.to_scipy()
    op = rcg.CsrOperator(A)
    rng = np.random.default_rng(6)
    b = 20
    X = rng.standard_normal((A.n, b)) + 1j * rng.standard_normal((A.n, b))
    # filter vs the oracle's complex filter
    ref = orc.filter_csr(A, X, 7, 8.5, 1.5)
    got = run_filter(rcg, op, X, 7, 8.5, 1.5)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    # CholQR: Q^H Q = I, R = Q^H X upper triangular with a positive real diagonal
    bx, bt = rcg.Block.from_host(X), rcg.Block(A.n, b, dtype=np.complex128)
    capi.check(capi.lib().rcg_orthonormalize(bx.h, bt.h, b))
    Q = bx.download()
    assert np.max(np.abs(Q.conj().T @ Q - np.eye(b))) < 1e-13
    R = Q.conj().T @ X
    assert np.max(np.abs(np.tril(R, -1))) < 1e-10 * np.max(np.abs(R))
    assert np.all(np.abs(np.diag(R).imag) < 1e-10 * np.abs(np.diag(R)))
    assert np.all(np.diag(R).real > 0)
    Qr, Rr = np.linalg.qr(X)
    ph = np.diag(Rr) / np.abs(np.diag(Rr))
    assert np.max(np.abs(Q - Qr * ph)) < 1e-10
    # deflation against 12 orthonormal complex vectors
    Y, _ = np.linalg.qr(rng.standard_normal((A.n, 12)) + 1j * rng.standard_normal((A.n, 12)))
    Z0 = rng.standard_normal((A.n, 8)) + 1j * rng.standard_normal((A.n, 8))
    bz, by = rcg.Block.from_host(Z0), rcg.Block.from_host(Y)
    capi.check(capi.lib().rcg_deflate(bz.h, 8, by.h, 12))
    Z = bz.download()
    P = Z0 - Y @ (Y.conj().T @ Z0)
    P /= np.linalg.norm(P, axis=0)
    assert np.max(np.abs(Y.conj().T @ Z)) < 1e-14
    assert np.max(np.abs(Z - P)) < 1e-13
    # projection H = Q^H A Q (complex, interleaved) and rotation
    bw = rcg.Block(A.n, b, dtype=np.complex128)
    H = np.zeros((b, b), dtype=np.complex128, order="F")
    capi.check(capi.lib().rcg_project(op.h, bx.h, bw.h, b, H.ctypes.data))
    Href = Q.conj().T @ (S @ Q)
    assert np.max(np.abs(H - Href)) < 1e-13 * np.max(np.abs(Href))
    assert np.array_equal(H, H.conj().T)
    w, Yh = np.linalg.eigh(H)
    Yh = np.asfortranarray(Yh)
    capi.check(capi.lib().rcg_rotate(bx.h, bt.h, b, Yh.ctypes.data))
    Vr = bx.download()
    ref = Q @ Yh
    ref /= np.linalg.norm(ref, axis=0)
    assert np.max(np.abs(Vr - ref)) < 1e-13


# ------------------------------------------------- benchmark-shaped solves
def test_anderson32_inc32_buf32_golden(rcg):
    """The 128^3 workload's solve shape (k 64, inc 32, buf 32, residual-only
    stop at 1e-10) on the 32^3 Anderson operator, against the compiled
    reference's golden record."""
    name = "anderson3d32_k64"
    c = CASES[name]
    print(f"case: {name} {c} vs the golden record", flush=True)
    A = build_operator(c)
    rc = rcg.RayleighChebyshevLM()
    rc.setStopCondition(c["stop"])
    op = rcg.CsrOperator(A, grid=(32, 32, 32))
    V = rcg.Block(A.n, 64)
    ev = []
    f = rc.getMinEigenSystem(c["k"], c["tol"], c["inc"], c["buf"], None, op,
                             rcg.RandomizeOp(c["seed"]), ev, V)
    g = GOLDEN[name]
    ev = np.array(ev)
    gev = np.array(g["eigenvalues"])
    assert f == g["found"] == 64
    assert rc.getSpectralBoundsIterCount() == g["lanczos_iters"]
    floor = np.maximum(np.abs(gev), 1e-11 / c["tol"])
    err = np.max(np.abs(ev - gev) / floor)
    print(f"max relative eigenvalue error vs the reference: {err:.2e}")
    assert err <= 1e-10
    X = V.download(f)
    normA = max(abs(rc.getMaxEigValueEst()), abs(rc.getMinEigValueEst()))
    r = np.linalg.norm(A.to_scipy() @ X - X * ev, axis=0) / normA
    assert np.max(r) <= c["tol"]


@pytest.mark.parametrize("mixed", [False, True])
def test_complex_k128_inc64_buf32(rcg, orc, mixed):
    """The complex workload's solve shape (k 128, inc 64, buf 32: 96 complex
    columns) on a 64 x 64 lattice at phi = 1/64 with onsite U[-1, 1]."""
    L, k, inc, buf, tol = 64, 128, 64, 32, 1e-10
    print(f"case: complex solve Hofstadter {L}x{L} phi=1/64 W=2 k={k} inc={inc} buf={buf} "
          f"mixed={mixed}", flush=True)
    A = ops.hofstadter_2d(L, phi=1 / 64, W=2.0, seed=2)
    rc = rcg.RayleighChebyshevLM()
    rc.setStopCondition("RESIDUAL_ONLY")
    if mixed:
        rc.filterPrecision = np.float32
    op = rcg.CsrOperator(A)
    V = rcg.Block(A.n, k, dtype=np.complex128)
    ev = []
    f = rc.getMinEigenSystem(k, tol, inc, buf, None, op, rcg.RandomizeOp(0), ev, V)
    ev = np.array(ev)
    o = orc.solve(A, k, tol, inc, buf, seed=0, stop="RESIDUAL_ONLY")
    assert f == o["found"] == k
    floor = np.maximum(np.abs(o["eigenvalues"]), 1e-11 / tol)
    err = np.max(np.abs(ev - o["eigenvalues"]) / floor)
    print(f"max relative eigenvalue error vs oracle: {err:.2e}")
    assert err <= 1e-10
    X = V.download(f)
    normA = max(abs(rc.getMaxEigValueEst()), abs(rc.getMinEigValueEst()))
    r = np.linalg.norm(A.to_scipy() @ X - X * ev, axis=0) / normA
    print(f"max ||Av - lambda v||/||A||: {np.max(r):.2e}")
    assert np.max(r) <= tol


def test_dense_filter_256_columns(rcg, orc):
    print("case: dense filter n=1500, 256 columns, degree 5 (FP64 and FP32 MFMA)", flush=True)
    n, b = 1500, 256
    lam = np.linspace(0.0, 1.0, n) ** 2
    G = np.random.default_rng(3).standard_normal((n, n))
    Q, R = np.linalg.qr(G)
    Q = Q * np.sign(np.diag(R))
    A = np.asfortranarray(0.5 * ((Q * lam) @ Q.T + ((Q * lam) @ Q.T).T))
    op = rcg.DenseOperator(A)
    X = np.ascontiguousarray(Q[:, :b])
    ref = orc.filter_dense(A, X, 5, 1.001, -lam[0])
    scale = np.max(np.abs(ref))
    bx, bt = rcg.Block.from_host(X), rcg.Block(n, b)
    capi.check(capi.lib().rcg_cheb_filter(op.h, bx.h, bt.h, b, 5, 1, 1.001, -lam[0]))
    assert np.max(np.abs(bx.download() - ref)) <= 1e-12 * scale
    th = np.ascontiguousarray(lam[:b])
    rz = np.ones(b, dtype=np.int32)
    bx = rcg.Block.from_host(X)
    capi.check(capi.lib().rcg_cheb_filter_mixed(op.h, bx.h, bt.h, b, 5, 1, 1.001, -lam[0],
                                                th.ctypes.data, rz.ctypes.data))
    assert np.max(np.abs(bx.download() - ref)) <= 1e-11 * scale


@pytest.mark.parametrize("mixed", [False, True])
def test_dense_inc128_buf128(rcg, mixed):
    """The dense workload's solve shape (inc 128, buf 128: 256 columns, k 256)
    at n = 2048 against the known spectrum linspace(0, 1, n)^2."""
    n, k, inc, buf, tol = 2048, 256, 128, 128, 1e-10
    print(f"case: dense solve n={n} k={k} inc={inc} buf={buf} mixed={mixed}", flush=True)
    lam = np.linspace(0.0, 1.0, n) ** 2
    G = np.random.default_rng(3).standard_normal((n, n))
    Q, R = np.linalg.qr(G)
    Q = Q * np.sign(np.diag(R))
    A = (Q * lam) @ Q.T
    A = np.asfortranarray(0.5 * (A + A.T))
    rc = rcg.RayleighChebyshevLM()
    rc.setStopCondition("RESIDUAL_ONLY")
    if mixed:
        rc.filterPrecision = np.float32
    op = rcg.DenseOperator(A)
    V = rcg.Block(n, inc + buf)
    ev = []
    f = rc.getMinEigenSystem(k, tol, inc, buf, None, op, rcg.RandomizeOp(0), ev, V)
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
