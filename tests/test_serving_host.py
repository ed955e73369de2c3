"""Host side of serving: the request stream (deterministic, distribution as
specified) and the footprint formula -- no GPU needed."""
import math

import numpy as np

from rayleighchebyshevgpgpu_amd import serving


def test_request_stream_is_deterministic_and_shaped():
    a = serving.serving_requests(4096)
    b = serving.serving_requests(4096)
    assert [(r.arrival, r.side, r.dim, r.W, r.k) for r in a] == [(r.arrival, r.side, r.dim, r.W, r.k) for r in b]
    t = np.array([r.arrival for r in a])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    # Poisson at 200/s: 4096 arrivals span about 20.5 s (mean gap 5 ms)
    assert 18.0 < t[-1] < 23.0
    assert abs(np.mean(gaps) - 1 / 200) < 3e-4
    n = np.array([r.n for r in a])
    assert n.min() >= 700 and n.max() <= 300_000
    # log-uniform N: the median is near sqrt(1e3 * 2.56e5) = 16000
    assert 10_000 < np.median(n) < 25_000
    ks = [r.k for r in a]
    assert set(ks) == {8, 16, 32} and all(abs(ks.count(k) / 4096 - 1 / 3) < 0.03 for k in (8, 16, 32))
    dims = [r.dim for r in a]
    assert abs(dims.count(2) / 4096 - 0.5) < 0.03
    W = np.array([r.W for r in a])
    assert W.min() >= 0 and W.max() <= 8 and abs(W.mean() - 4) < 0.15


def test_request_operator_shape():
    r = serving.ServingRequest(0, 0.0, 3, 10, 2.0, 8, 1e-10, 1000)
    A = r.operator()
    assert A.n == 1000 and A.nnz == 1000 + 2 * 3 * 10 * 10 * 9
    d = A.to_scipy().diagonal()
    assert np.all(np.abs(d - 6.0) <= 1.0)


def test_footprint_formula():
    fp = serving.solve_footprint(10_000, 16, 16, 8)
    ld = 24
    assert fp == 8 * 10_000 * (3 * ld + 2 * ld) + 4 * 3 * 10_000 * ld + 8 * 2 * 10_000 * ld
    assert serving.solve_footprint(10_000, 16, 16, 8, mixed=False) < fp
    assert serving.solve_footprint(10_000, 16, 16, 8, is_complex=True) > fp
