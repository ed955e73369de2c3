"""Cases recorded from the compiled reference into ref_cases.json.

Each case names an operator (built by rayleighchebyshevgpgpu_amd.operators)
and the getMinEigenSystem / getMinIntervalEigenSystem arguments.
"""
from __future__ import annotations

from rayleighchebyshevgpgpu_amd import operators as ops

CASES = {
    # 2-D 5-point Dirichlet Laplacian 64x64, lowest 16, inc 8, buf 8, tol 1e-8, seed 0
    "lap2d64_k16": dict(op="lap2d", nx=64, bc="dirichlet", k=16, inc=8, buf=8, tol=1e-8, seed=0),
    # same operator, residual-only stop at 1e-10
    "lap2d64_k16_resid": dict(op="lap2d", nx=64, bc="dirichlet", k=16, inc=8, buf=8, tol=1e-10,
                              seed=0, stop="RESIDUAL_ONLY"),
    # eigenvalue-only stop
    "lap2d40_k12_eig": dict(op="lap2d", nx=40, bc="dirichlet", k=12, inc=6, buf=6, tol=1e-9,
                            seed=5, stop="EIGENVALUE_ONLY"),
    # 3-D periodic 16^3 + Anderson W=4 (seed 1): a reduced 128^3 Anderson operator
    "anderson3d16_k32": dict(op="lap3d", nx=16, bc="periodic", W=4.0, pseed=1, k=32, inc=16,
                             buf=16, tol=1e-10, seed=0, stop="RESIDUAL_ONLY"),
    # 3-D periodic 32^3 + Anderson W=4: inc 32, buf 32 as in the 128^3 workload
    "anderson3d32_k64": dict(op="lap3d", nx=32, bc="periodic", W=4.0, pseed=1, k=64, inc=32,
                             buf=32, tol=1e-10, seed=0, stop="RESIDUAL_ONLY"),
    # 2-D Dirichlet 48x48 + potential W=8 (seed 4), small buffer
    "anderson2d48_k16": dict(op="lap2d", nx=48, bc="dirichlet", W=8.0, pseed=4, k=16, inc=8,
                             buf=4, tol=1e-10, seed=7),
    # near-full basis: N=36, 30 eigenpairs, exercises the subspace shrink / completion path
    "lap2d6_k30_fill": dict(op="lap2d", nx=6, bc="dirichlet", k=30, inc=8, buf=8, tol=1e-10,
                            seed=3),
    # subspace >= N: quick-return path (N=16)
    "lap2d4_quick": dict(op="lap2d", nx=4, bc="dirichlet", k=4, inc=8, buf=8, tol=1e-10, seed=0),
    # every eigenpair below 0.1 (interval variant), at most 100
    "lap2d32_interval": dict(op="lap2d", nx=32, bc="dirichlet", k=100, inc=8, buf=8, tol=1e-10,
                             seed=2, interval=True, lambda_max=0.1),
    # iteration cap of 1: unconverged return
    "lap2d64_maxinner1": dict(op="lap2d", nx=64, bc="dirichlet", k=16, inc=8, buf=8, tol=1e-8,
                              seed=0, max_inner=1),
}


def build_operator(c):
    if c["op"] == "lap2d":
        n = c["nx"] * c["nx"]
        pot = ops.anderson_potential(n, c["W"], c["pseed"]) if c.get("W") else None
        return ops.laplacian_2d(c["nx"], bc=c["bc"], potential=pot)
    if c["op"] == "lap3d":
        n = c["nx"] ** 3
        pot = ops.anderson_potential(n, c["W"], c["pseed"]) if c.get("W") else None
        return ops.laplacian_3d(c["nx"], bc=c["bc"], potential=pot)
    raise ValueError(c["op"])
