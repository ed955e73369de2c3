"""The CPU oracle against the compiled reference (tests/golden/ref_cases.json).

The fixture holds what the reference's own RayleighChebyshevLM returned when
built from its headers (oracle/Makefile `ref`) on each case of
tests/golden/cases.py.  The restatement must reproduce it: the same number of
eigenpairs, Lanczos iterations, random draws, countData and filter-degree
sequence, and eigenvalues to 1e-12 relative (floor max(|lambda|, 1e-11/tol),
RayleighChebyshevLM.h:271-278).
"""
import json
import os

import numpy as np
import pytest

from tests.golden.cases import CASES, build_operator

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_cases.json")))

# The reference returns stale (unsorted) eigenvalues and, when foundSize differs
# from the subspace size, residuals of a wrongly formed projection after its
# non-monotone repair (RayleighChebyshevLM.h:1664-1707, 1932).  The oracle
# stores the recomputed Ritz values; for these cases only the multiset of
# eigenvalues before the repair, the trajectory and counters are compared.
REPAIRED = {"lap2d6_k30_fill", "lap2d64_maxinner1"}
# The residual-only stop of these cases ends through the oscillation guard
# (RayleighChebyshevLM.h:1418-1445), which compares second differences of
# residuals at the 1e-14 level: the tail length is roundoff-sensitive.
OSCILLATION_TAIL = {"anderson3d32_k64"}

FAST = [n for n in CASES if n != "anderson3d32_k64"]


def _run(orc, name):
    c = CASES[name]
    A = build_operator(c)
    return c, orc.solve(A, c["k"], c["tol"], c["inc"], c["buf"], seed=c.get("seed", 0),
                        stop=c.get("stop", "COMBINATION"), interval=c.get("interval", False),
                        lambda_max=c.get("lambda_max", 0.0), max_inner=c.get("max_inner", 0))


def _check(orc, name):
    c, r = _run(orc, name)
    g = GOLDEN[name]
    assert r["found"] == g["found"]
    assert r["lanczos_iters"] == g["lanczos_iters"]
    assert abs(r["max_est"] - g["max_est"]) <= 1e-12 * abs(g["max_est"])
    assert abs(r["min_est"] - g["min_est"]) <= 1e-10 * max(abs(g["min_est"]), 1.0)
    ref_deg = [int(t[1]) for t in g["trace"]]
    our_deg = r["trace"][:, 2].astype(int).tolist()
    if name in OSCILLATION_TAIL:
        m = min(len(ref_deg), len(our_deg))
        assert our_deg[:m] == ref_deg[:m]
        assert abs(len(ref_deg) - len(our_deg)) <= 5
    else:
        assert our_deg == ref_deg
        assert r["count"] == g["count"]
        assert r["rng_counter"] == g["rng_counter"]
    ge = np.array(g["eigenvalues"])
    floor = np.maximum(np.abs(ge), 1e-11 / max(c["tol"], 1e-11))
    if name in REPAIRED:
        # multiset of Ritz values: compare sorted lists to the repair's tolerance
        assert np.all(np.abs(np.sort(ge) - np.sort(r["eigenvalues"])) <= 2.0)
        return
    assert np.all(np.abs(r["eigenvalues"] - ge) / floor <= 1e-12)
    gr = np.array(g["residuals"])
    assert np.all(np.abs(r["residuals"] - gr) <= 1e-9 * np.maximum(1.0, np.abs(ge)) + 1e-3 * gr)


@pytest.mark.parametrize("name", FAST)
def test_oracle_matches_compiled_reference(orc, name):
    _check(orc, name)


@pytest.mark.slow
def test_oracle_matches_compiled_reference_32cube(orc):
    _check(orc, "anderson3d32_k64")


def test_repaired_cases_are_sorted_and_accurate(orc):
    """Where the reference leaves stale values, the oracle's are ascending and
    satisfy their residuals (checked against the dense spectrum)."""
    from rayleighchebyshevgpgpu_amd import operators as ops
    c, r = _run(orc, "lap2d6_k30_fill")
    assert np.all(np.diff(r["eigenvalues"]) >= -1e-12)
    exact = ops.laplacian_2d_spectrum(6)[:30]
    assert np.max(np.abs(r["eigenvalues"] - exact)) < 1e-9
    assert np.max(r["residuals"]) < np.sqrt(c["tol"])  # COMBINATION stop: residual <= sqrt(tol)
