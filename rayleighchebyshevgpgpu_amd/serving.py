"""Serving: a stream of lowest-eigenpair requests through the continuous-
batching solve server of librcg.so (rcg_server_*, include/rcg_c_api.h).

The reference solves one operator per call; this module is the request
side of the batched workload: ``ServingRequest`` streams generated from one
seeded counter-based stream, and ``Server``, a thin ctypes wrapper that
submits requests, waits for tickets and reads the timings, the results and
the eigenvectors left resident on the device.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import capi
from . import operators as ops
from .capi import check, lib
from .solver import Context, default_context


@dataclass
class ServingRequest:
    index: int
    arrival: float          # seconds after the stream starts
    dim: int                # 2 or 3
    side: int               # grid points per side (Dirichlet boundary)
    W: float                # potential width: onsite U[-W/2, W/2]
    k: int                  # eigenpairs wanted
    tol: float = 1e-10
    seed: int = 0           # potential stream (derived from the request index)

    @property
    def n(self) -> int:
        return self.side ** self.dim

    def potential(self):
        return ops.anderson_potential(self.n, self.W, self.seed) if self.W > 0 else None

    def operator(self) -> ops.Csr:
        """The request's 5- or 7-point Dirichlet Laplacian plus its potential."""
        if self.dim == 2:
            return ops.laplacian_2d(self.side, bc="dirichlet", potential=self.potential())
        return ops.laplacian_3d(self.side, bc="dirichlet", potential=self.potential())


class OperatorCache:
    """The stream's operators on the device, one sparsity pattern per grid:
    the first request of a (dimension, side) uploads its CSR; later ones share
    that pattern's row pointers and columns (rcg_operator_csr_share) and upload
    only their values -- the same arrays as ``ServingRequest.operator()``."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx
        self.base = {}

    def operator(self, r: ServingRequest):
        from .solver import CsrOperator
        key = (r.dim, r.side)
        if key not in self.base:
            A0 = (ops.laplacian_2d(r.side, bc="dirichlet") if r.dim == 2
                  else ops.laplacian_3d(r.side, bc="dirichlet"))
            rows = np.repeat(np.arange(A0.n), np.diff(A0.rowptr))
            diag = np.flatnonzero(A0.col == rows)
            self.base[key] = (CsrOperator(A0, self.ctx), A0.val, diag)
        op0, val0, diag = self.base[key]
        val = val0.copy()
        pot = r.potential()
        if pot is not None:
            val[diag] = val0[diag] + pot  # 2d + V_i, as the host builder forms it
        return op0.share(val)


def serving_requests(count: int = 4096, rate: float = 200.0, seed: int = 4,
                     n_min: float = 1e3, n_max: float = 2.56e5) -> list[ServingRequest]:
    """The batched-serving request stream: Poisson arrivals at `rate` per
    second; N log-uniform in [n_min, n_max] snapped to a 2-D (side round(sqrt N))
    or 3-D (side round(cbrt N)) grid by a fair coin; W uniform in [0, 8];
    k uniform in {8, 16, 32}; tolerance 1e-10.  Request i takes the six draws
    6i .. 6i+5 of stream `seed` (inter-arrival, N, coin, W, k, spare) and its
    potential comes from stream 1000 + i."""
    u = ops.unit_draws(seed, 0, 6 * count).reshape(count, 6)
    t = 0.0
    out = []
    for i in range(count):
        t += -math.log(1.0 - u[i, 0]) / rate
        N = math.exp(math.log(n_min) + (math.log(n_max) - math.log(n_min)) * u[i, 1])
        dim = 2 if u[i, 2] < 0.5 else 3
        side = max(4, int(round(N ** (1.0 / dim))))
        W = 8.0 * u[i, 3]
        k = (8, 16, 32)[min(2, int(3 * u[i, 4]))]
        out.append(ServingRequest(i, t, dim, side, W, k, 1e-10, 1000 + i))
    return out


def _block_ld(cols: int) -> int:
    return max(4, (cols + 3) // 4 * 4)


def solve_footprint(n: int, k: int, inc: int, buf: int, is_complex: bool = False,
                    mixed: bool = True, refine: bool = False) -> int:
    """Device bytes the server charges a request against its HBM budget: the
    solve's V, T, W and found-vector blocks plus the caller's copy of the
    found vectors, the FP32 correction blocks of the mixed filter and the
    refinement blocks."""
    cw = 2 if is_complex else 1
    sub = min(inc + buf, n)
    ld = cw * _block_ld(sub)
    ldy = cw * _block_ld(max(min(k, n), sub))
    b = 8 * n * (3 * ld + 2 * ldy)
    if mixed:
        b += 4 * 3 * n * ld
    if refine or mixed:
        b += 8 * 2 * n * ldy
    return b


class Server:
    """rcg_server: admission by HBM footprint, shared filter batches."""

    def __init__(self, ctx: Context | None = None, hbm_budget: int = 64 << 30,
                 max_active: int = 32, window_us: int = 200):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        check(lib().rcg_server_create(self.ctx.h, int(hbm_budget), int(max_active),
                                      int(window_us), C.byref(h)))
        self.h = h

    def submit(self, op, k: int, tol: float, inc: int, buf: int, precision: str = "mixed",
               seed: int = 0, stop: str = "RESIDUAL_ONLY", refine: bool = False,
               start_ticket: int | None = None) -> int:
        """Queue a solve; with start_ticket, it starts from that finished
        ticket's resident eigenvectors (warm start)."""
        p = capi.SolveParams(k=int(k), inc=int(inc), buf=int(buf), tol=float(tol), seed=int(seed),
                             stop=capi.STOP[stop], filter_dtype=capi.F32 if precision == "mixed"
                             else capi.F64, refine=int(refine))
        t = C.c_int64()
        if start_ticket is None:
            check(lib().rcg_server_submit(self.h, op.h, C.byref(p), C.byref(t)))
        else:
            check(lib().rcg_server_submit_warm(self.h, op.h, C.byref(p), int(start_ticket),
                                               C.byref(t)))
        return t.value

    def wait(self, ticket: int, cap: int = 4096) -> dict:
        r = capi.SolveResult()
        ev = np.zeros(cap)
        rr = np.zeros(cap)
        tm = np.zeros(3)
        code = lib().rcg_server_wait(self.h, ticket, C.byref(r), ev.ctypes.data_as(C.c_void_p),
                                     rr.ctypes.data_as(C.c_void_p), cap,
                                     tm.ctypes.data_as(C.c_void_p))
        f = int(r.found)
        return {"status": code, "error": capi.last_error() if code else "", "found": f,
                "eigenvalues": ev[:f].copy(), "residuals": rr[:f].copy(), "result": r,
                "t_submit": tm[0], "t_admit": tm[1], "t_done": tm[2]}

    def vectors(self, ticket: int, n: int, cols: int, dtype=np.float64) -> np.ndarray:
        """Download the ticket's resident eigenvectors (n x cols)."""
        b = C.c_void_p()
        check(lib().rcg_server_vectors(self.h, ticket, C.byref(b)))
        X = np.empty((n, cols), dtype=dtype, order="F")
        check(lib().rcg_block_download(b, X.ctypes.data_as(C.c_void_p), n, cols))
        return X

    def release(self, ticket: int):
        check(lib().rcg_server_release(self.h, ticket))

    def stats(self) -> dict:
        s = np.zeros(6, dtype=np.int64)
        check(lib().rcg_server_stats(self.h, s.ctypes.data_as(C.c_void_p)))
        return {"completed": int(s[0]), "filter_batches": int(s[1]), "batched_requests": int(s[2]),
                "hbm_in_use": int(s[3]), "hbm_peak": int(s[4]), "queued": int(s[5])}

    def filter_profile(self) -> dict:
        """Device time, launches and algorithmic bytes of the shared filter launches."""
        o = np.zeros(4)
        check(lib().rcg_server_filter_profile(self.h, o.ctypes.data_as(C.c_void_p)))
        return {"device_s": float(o[0]), "launches": int(o[1]), "alg_bytes": float(o[2]),
                "batches_timed": int(o[3])}

    def close(self):
        if getattr(self, "h", None):
            lib().rcg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
