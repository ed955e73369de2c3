"""Builds the dense workload's operator on the GPU with torch, in a process of
its own, and writes it to a .npy file:

    python -m rayleighchebyshevgpgpu_amd.dense_gen N SEED OUT.npy

A = Q diag(lambda) Q^T with lambda = linspace(0, 1, N)^2 and Q = qr(G).Q
diag(sign(diag R)), G ~ N(0, 1) drawn column by column by Box-Muller from the
counter stream SEED -- the same matrix as operators.dense_qlq, which does
the QR on the host and is too slow at N = 16384.  torch is only the input
generator here (its own HIP runtime, in this child process); the solver
reads the file.
"""
from __future__ import annotations

import sys
import time

import numpy as np


def gaussian_columns(n: int, seed: int, c0: int, c1: int) -> np.ndarray:
    from . import operators as ops
    cnt = (c1 - c0) * n
    u = ops.unit_draws(seed, c0 * n, cnt + cnt % 2)
    u1, u2 = 1.0 - u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.empty(u.shape[0])
    g[0::2] = r * np.cos(2 * np.pi * u2)
    g[1::2] = r * np.sin(2 * np.pi * u2)
    return g[:cnt].reshape(c1 - c0, n).T  # n x (c1 - c0), column j = draws of column c0 + j


def main() -> int:
    n, seed, out = int(sys.argv[1]), int(sys.argv[2]), sys.argv[3]
    import torch
    if n % 2 != 0:
        raise SystemExit("N must be even (Box-Muller pairs within a column chunk)")
    dev = torch.device("cuda:0")
    t0 = time.time()
    G = torch.empty((n, n), dtype=torch.float64, device=dev)
    step = max(2, (1 << 25) // n)
    for c0 in range(0, n, step):
        c1 = min(n, c0 + step)
        G[:, c0:c1] = torch.from_numpy(np.ascontiguousarray(gaussian_columns(n, seed, c0, c1))).to(dev)
    print(f"[dense_gen] G drawn in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    Q, R = torch.linalg.qr(G)
    del G
    Q = Q * torch.sign(torch.diagonal(R))[None, :]
    del R
    torch.cuda.synchronize()
    print(f"[dense_gen] QR done at {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    lam = torch.linspace(0.0, 1.0, n, dtype=torch.float64, device=dev) ** 2
    A = (Q * lam[None, :]) @ Q.T
    A = 0.5 * (A + A.T)
    np.save(out, A.cpu().numpy())
    print(f"[dense_gen] A written in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
