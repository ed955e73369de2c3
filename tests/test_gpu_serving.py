"""The continuous-batching solve server (rcg_server_*).

Pins: each request's result is the same alone (one worker, nothing to batch
with) and in the mix (many workers, shared filter batches) -- the batched
step computes every row exactly as the single-operator step does; served
eigenvalues match the analytic Dirichlet spectra and the FP64 oracle to a
relative 1e-10; every pair's residual meets the tolerance; the HBM budget is
never exceeded by the admitted solves; a failing request reports its own
error without stopping the others.
"""
import numpy as np
import pytest

from rayleighchebyshevgpgpu_amd import operators as ops
from rayleighchebyshevgpgpu_amd import serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rcg():
    from rayleighchebyshevgpgpu_amd import solver
    return solver


def small_requests(count, seed=11):
    reqs = serving.serving_requests(count, rate=1e6, seed=seed, n_min=300, n_max=6000)
    return reqs


def submit_all(srv, rcg, reqs, precision):
    ops_, tickets = [], []
    for r in reqs:
        op = rcg.CsrOperator(r.operator())
        ops_.append(op)
        tickets.append(srv.submit(op, r.k, r.tol, r.k, max(4, r.k // 2), precision=precision))
    return ops_, tickets


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_alone_equals_mix(rcg, precision):
    reqs = small_requests(10)
    print(f"case: 10 requests alone vs in the mix, {precision}", flush=True)
    alone = []
    srv1 = serving.Server(max_active=1, window_us=0)
    for r in reqs:
        op = rcg.CsrOperator(r.operator())
        t = srv1.submit(op, r.k, r.tol, r.k, max(4, r.k // 2), precision=precision)
        out = srv1.wait(t)
        assert out["status"] == 0, out["error"]
        alone.append(out)
        srv1.release(t)
    srv1.close()
    srv = serving.Server(max_active=8, window_us=300)
    ops_, tickets = submit_all(srv, rcg, reqs, precision)
    mix = [srv.wait(t) for t in tickets]
    st = srv.stats()
    print(f"server stats: {st}")
    assert st["batched_requests"] > st["filter_batches"]  # some batches held several solves
    for r, a, m in zip(reqs, alone, mix):
        assert m["status"] == 0, m["error"]
        assert a["found"] == m["found"] == r.k
        d = np.max(np.abs(a["eigenvalues"] - m["eigenvalues"]) / np.maximum(np.abs(a["eigenvalues"]), 0.1))
        assert d <= 1e-12, (r, d)
    srv.close()


@pytest.mark.skip(reason="the persistent one-launch sweep of a batch (tasks ordered by step, "
                         "workgroups waiting on per-request completion counters) stalled the "
                         "device on its first served request in two runs; it is not in the "
                         "library, and RCG_BATCH_SWEEP selects nothing until it is rebuilt")
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_sweep_equals_per_step_launches(rcg, precision, monkeypatch):
    """The persistent sweep (every step of a batch in one launch, tasks ordered
    by step with per-request completion counters) computes each row exactly as
    the per-step batched launches do: the served results agree as alone and in
    the mix (1e-12)."""
    reqs = small_requests(8, seed=21)
    print(f"case: 8 served requests, persistent sweep vs per-step launches, {precision}", flush=True)
    outs = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("RCG_BATCH_SWEEP", mode)
        srv = serving.Server(max_active=8, window_us=300)
        ops_, tickets = submit_all(srv, rcg, reqs, precision)
        outs[mode] = [srv.wait(t) for t in tickets]
        st = srv.stats()
        assert st["batched_requests"] > st["filter_batches"]
        srv.close()
    for r, a, b in zip(reqs, outs["1"], outs["0"]):
        assert a["status"] == 0 and b["status"] == 0, (a["error"], b["error"])
        assert a["found"] == b["found"] == r.k
        d = np.max(np.abs(a["eigenvalues"] - b["eigenvalues"]) / np.maximum(np.abs(a["eigenvalues"]), 0.1))
        assert d <= 1e-12, (r, d)


def test_served_results_vs_analytic_and_oracle(rcg, orc):
    print("case: served W=0 Dirichlet Laplacians vs closed forms and the oracle", flush=True)
    shapes = [(2, 40, 16), (3, 14, 8), (2, 64, 32), (3, 10, 16)]
    srv = serving.Server(max_active=4, window_us=300)
    held, tickets = [], []
    for dim, side, k in shapes:
        A = ops.laplacian_2d(side) if dim == 2 else ops.laplacian_3d(side, bc="dirichlet")
        op = rcg.CsrOperator(A)
        held.append((A, op, dim, side, k))
        tickets.append(srv.submit(op, k, 1e-10, k, k // 2, precision="mixed"))
    for (A, op, dim, side, k), t in zip(held, tickets):
        out = srv.wait(t)
        assert out["status"] == 0, out["error"]
        ev = out["eigenvalues"]
        exact = (ops.laplacian_2d_spectrum(side) if dim == 2
                 else ops.laplacian_3d_spectrum(side, bc="dirichlet"))[:k]
        assert np.max(np.abs(ev - exact) / np.maximum(exact, 0.1)) < 1e-10
        o = orc.solve(A, k, 1e-10, k, k // 2, seed=0, stop="RESIDUAL_ONLY")
        assert np.max(np.abs(ev - o["eigenvalues"]) / np.maximum(np.abs(ev), 0.1)) < 1e-10
        X = srv.vectors(t, A.n, k)
        r = np.linalg.norm(A.to_scipy() @ X - X * ev, axis=0) / (4.0 * dim)
        assert np.max(r) <= 1e-10
        srv.release(t)
    srv.close()


def test_hbm_budget_respected(rcg):
    reqs = small_requests(12, seed=12)
    print("case: HBM budget of about three requests, 8 worker slots", flush=True)
    # a budget of 6 MB admits a few of these at a time
    srv = serving.Server(max_active=8, hbm_budget=6 << 20, window_us=200)
    ops_, tickets = submit_all(srv, rcg, reqs, "mixed")
    outs = []
    for t in tickets:
        out = srv.wait(t)
        assert out["status"] == 0, out["error"]
        assert out["t_submit"] <= out["t_admit"] <= out["t_done"]
        outs.append(out)
    st = srv.stats()
    print(f"server stats, all done, none released: {st}")
    # finished, unreleased tickets keep their eigenvectors on the device and
    # count against the budget: at least n x ld(found) FP64 each
    resident = sum(r.n * serving._block_ld(o["found"]) * 8 for r, o in zip(reqs, outs))
    assert st["hbm_in_use"] >= resident and st["queued"] == 0
    fps = [serving.solve_footprint(r.n, r.k, r.k, max(4, r.k // 2)) for r in reqs]
    # a request larger than the whole budget is admitted only when no other
    # solve runs, next to at most the resident vectors of finished tickets
    assert st["hbm_peak"] <= max(6 << 20, max(fps) + st["hbm_in_use"])
    assert st["hbm_peak"] >= max(fps)
    for t in tickets:
        srv.release(t)
    st = srv.stats()
    assert st["hbm_in_use"] == 0 and st["queued"] == 0
    srv.close()


def test_resident_vectors_hold_back_admission(rcg):
    """Unreleased results count against the budget until their release; a
    request that does not fit next to them is still admitted once no other
    solve runs, so resident vectors cannot hold the queue back for good."""
    print("case: resident vectors of unreleased tickets against a small budget", flush=True)
    A = ops.laplacian_2d(40)
    k = 8
    fp = serving.solve_footprint(A.n, k, k, 4)
    op = rcg.CsrOperator(A)
    # room for one running solve plus a little: two unreleased results and a
    # running solve do not fit together
    budget = fp + 4096
    srv = serving.Server(max_active=4, hbm_budget=budget, window_us=0)
    t1 = srv.submit(op, k, 1e-10, k, 4, precision="fp64")
    r1 = srv.wait(t1)
    assert r1["status"] == 0, r1["error"]
    held = srv.stats()["hbm_in_use"]
    assert held >= A.n * serving._block_ld(k) * 8
    # nothing runs, so the next request is admitted next to the resident block
    t2 = srv.submit(op, k, 1e-10, k, 4, precision="fp64")
    r2 = srv.wait(t2)
    assert r2["status"] == 0, r2["error"]
    st = srv.stats()
    assert st["hbm_peak"] <= fp + 2 * held
    assert st["hbm_in_use"] >= 2 * A.n * serving._block_ld(k) * 8
    srv.release(t1)
    srv.release(t2)
    assert srv.stats()["hbm_in_use"] == 0
    assert np.array_equal(r1["eigenvalues"], r2["eigenvalues"])
    srv.close()


def test_failing_request_does_not_stop_others(rcg):
    print("case: a non-Hermitian request fails alone; its neighbours complete", flush=True)
    n = 300
    rows = np.arange(n)
    bad = ops._from_coo(n, np.concatenate([rows, rows[:-1]]), np.concatenate([rows, rows[1:]]),
                        np.concatenate([np.ones(n), 1j * np.ones(n - 1)]).astype(complex))
    good = ops.laplacian_2d(30)
    srv = serving.Server(max_active=3, window_us=100)
    ob, og1, og2 = rcg.CsrOperator(bad), rcg.CsrOperator(good), rcg.CsrOperator(good)
    t1 = srv.submit(og1, 8, 1e-10, 8, 4)
    t2 = srv.submit(ob, 4, 1e-8, 4, 4)
    t3 = srv.submit(og2, 8, 1e-10, 8, 4)
    r1, r2, r3 = srv.wait(t1), srv.wait(t2), srv.wait(t3)
    assert r1["status"] == 0 and r3["status"] == 0
    assert r2["status"] == 4 and "Hermitian" in r2["error"]
    assert np.array_equal(r1["eigenvalues"], r3["eigenvalues"])
    srv.close()


def test_warm_start_from_resident_vectors(rcg):
    """A request started from a finished ticket's resident eigenvectors
    (setNonRandomStartFlag across requests): the same eigenvalues in fewer
    inner iterations, also for a nearby operator; the start ticket cannot be
    released while the warm solve is queued or running."""
    print("case: warm start from a finished ticket, 48x48 Dirichlet + potential", flush=True)
    side, k = 48, 12
    A0 = ops.laplacian_2d(side, potential=ops.anderson_potential(side * side, 1.0, 7))
    A1 = ops.laplacian_2d(side, potential=1.001 * ops.anderson_potential(side * side, 1.0, 7))
    op0, op1 = rcg.CsrOperator(A0), rcg.CsrOperator(A1)
    srv = serving.Server(max_active=2, window_us=0)
    cold = srv.wait(srv.submit(op0, k, 1e-10, k, 6, precision="fp64"))
    assert cold["status"] == 0 and cold["found"] == k
    t0 = srv.submit(op0, k, 1e-10, k, 6, precision="fp64")
    first = srv.wait(t0)
    tw = srv.submit(op0, k, 1e-10, k, 6, precision="fp64", start_ticket=t0)
    tn = srv.submit(op1, k, 1e-10, k, 6, precision="fp64", start_ticket=t0)
    warm, near = srv.wait(tw), srv.wait(tn)
    assert warm["status"] == 0 and near["status"] == 0
    floor = np.maximum(np.abs(first["eigenvalues"]), 0.1)  # the reference's relative floor
    assert np.max(np.abs(warm["eigenvalues"] - first["eigenvalues"]) / floor) < 1e-10
    assert warm["result"].count_eig < first["result"].count_eig
    ref1 = np.linalg.eigvalsh(A1.to_scipy().toarray())[:k]
    assert np.max(np.abs(near["eigenvalues"] - ref1) / np.maximum(np.abs(ref1), 0.1)) < 1e-10
    assert near["result"].count_eig < first["result"].count_eig
    srv.release(tw)
    srv.release(tn)
    srv.release(t0)
    srv.close()


def test_warm_start_rejects_bad_tickets(rcg):
    print("case: warm start argument checks", flush=True)
    A = ops.laplacian_2d(20)
    B = ops.laplacian_2d(21)
    srv = serving.Server(max_active=1, window_us=0)
    opa, opb = rcg.CsrOperator(A), rcg.CsrOperator(B)
    t = srv.submit(opa, 4, 1e-10, 4, 4, precision="fp64")
    srv.wait(t)
    with pytest.raises(Exception, match="do not match"):
        srv.submit(opb, 4, 1e-10, 4, 4, precision="fp64", start_ticket=t)
    with pytest.raises(Exception, match="unknown start ticket"):
        srv.submit(opa, 4, 1e-10, 4, 4, precision="fp64", start_ticket=999)
    srv.release(t)
    srv.close()


def test_shared_pattern_operators_match_host(rcg):
    """The serving bench uploads one sparsity pattern per grid and each
    request's values (rcg_operator_csr_share): every product equals the one of
    the request's own host-built operator, bit for bit."""
    print("case: shared-pattern serving operators vs their own CSR", flush=True)
    from rayleighchebyshevgpgpu_amd import capi
    reqs = serving.serving_requests(60, rate=200.0, seed=4, n_min=300, n_max=9000)
    # requests on repeated grids, 2-D and 3-D
    reqs = reqs + [serving.ServingRequest(100 + i, 0.0, r.dim, r.side, 3.0 + i, r.k, seed=2000 + i)
                   for i, r in enumerate(reqs[:4])]
    cache = serving.OperatorCache()
    rng = np.random.default_rng(0)
    shared = 0
    for r in reqs:
        op = cache.operator(r)
        ref = rcg.CsrOperator(r.operator())
        assert op.nnz == ref.nnz
        X = rng.standard_normal((r.n, 5))
        bx, b1, b2 = rcg.Block.from_host(X), rcg.Block(r.n, 5), rcg.Block(r.n, 5)
        op.apply(bx, b1)
        ref.apply(bx, b2)
        assert np.array_equal(b1.download(), b2.download())
        shared += (r.dim, r.side) in cache.base and op is not cache.base[(r.dim, r.side)][0]
    assert len(cache.base) < len(reqs) and shared == len(reqs)
    # a pattern from a non-CSR operator is refused
    D = rcg.DenseOperator(np.eye(4))
    h = capi.C.c_void_p()
    assert capi.lib().rcg_operator_csr_share(D.h, np.ones(16).ctypes.data, 0, capi.C.byref(h)) == 1
