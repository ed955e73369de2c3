"""Synthetic operators and the start-vector random stream.

The generators build the CSR (int32 indices, sorted columns) or dense
operators the solver is benchmarked and tested on:

* ``laplacian_2d`` / ``laplacian_3d``: 5-/7-point Laplacians (diagonal 2d,
  neighbours -1) with Dirichlet or periodic boundaries, lexicographic order
  with x fastest, plus an optional diagonal potential (Anderson disorder).
* ``hofstadter_2d``: the complex-Hermitian tight-binding Hamiltonian with a
  Peierls phase on y-bonds (Landau gauge).
* ``dense_qlq``: A = Q diag(lam) Q^T with a Haar-random orthogonal Q.

``random_stream`` is the counter-based SplitMix64 stream shared with the
device kernels (start vectors are generated in place on the GPU) and with the
CPU oracle: draw ``idx`` of seed ``s`` is ``mix64(mix64(s+g) + (idx+1) g)``
mapped to [0,1) by its top 53 bits.  Start vectors follow the consumption
order of ArandomizeOpClassLMCuda.h:58-91 (one draw per real element, two per
complex element, vectors in call order).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

__all__ = [
    "mix64", "unit_draws", "sym_draws", "uniform_draws",
    "laplacian_2d", "laplacian_3d", "hofstadter_2d", "dense_qlq",
    "laplacian_2d_spectrum", "laplacian_3d_spectrum", "Csr",
]

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def mix64(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def unit_draws(seed: int, start: int, count: int) -> np.ndarray:
    """Draws start..start+count-1 of stream `seed`, uniform in [0,1)."""
    with np.errstate(over="ignore"):
        key = mix64(np.uint64(seed) + _GAMMA)
        idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = mix64(key + idx * _GAMMA)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def sym_draws(seed: int, start: int, count: int) -> np.ndarray:
    """U[-1,1) draws, the start-vector distribution (exact: -1 + 2u)."""
    return -1.0 + 2.0 * unit_draws(seed, start, count)


def uniform_draws(seed: int, start: int, count: int, a: float, b: float) -> np.ndarray:
    return a + (b - a) * unit_draws(seed, start, count)


class Csr:
    """CSR operator with int32 indices (host arrays)."""

    def __init__(self, n, rowptr, col, val):
        self.n = int(n)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val)
        if self.val.dtype not in (np.float64, np.complex128):
            self.val = self.val.astype(np.float64)

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])

    @property
    def is_complex(self) -> bool:
        return self.val.dtype == np.complex128

    def to_scipy(self) -> sp.csr_matrix:
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n, self.n))

    def write_bin(self, path: str) -> None:
        """int64 n, int64 nnz, int32 rowptr, int32 col, float64 val."""
        with open(path, "wb") as f:
            np.array([self.n, self.nnz], dtype=np.int64).tofile(f)
            self.rowptr.tofile(f)
            self.col.tofile(f)
            self.val.astype(np.float64).tofile(f)


def _from_coo(n, rows, cols, vals) -> Csr:
    m = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    m.sum_duplicates()
    m.sort_indices()
    return Csr(n, m.indptr, m.indices, m.data)


def _grid_neighbours(shape, periodic):
    """(rows, cols) of nearest-neighbour bonds in both directions."""
    dims = len(shape)
    n = int(np.prod(shape))
    idx = np.arange(n, dtype=np.int64)
    coords = []
    rem = idx.copy()
    for d in range(dims):  # x fastest
        coords.append(rem % shape[d])
        rem //= shape[d]
    strides = [int(np.prod(shape[:d])) for d in range(dims)]
    rows, cols = [], []
    for d in range(dims):
        L = shape[d]
        c = coords[d]
        for step in (-1, 1):
            t = c + step
            if periodic:
                ok = np.ones(n, dtype=bool)
                t = t % L
            else:
                ok = (t >= 0) & (t < L)
            j = idx + (t - c) * strides[d]
            rows.append(idx[ok])
            cols.append(j[ok])
    return np.concatenate(rows), np.concatenate(cols)


def _laplacian(shape, bc, potential):
    """CSR of the 2d+1-point Laplacian on `shape` (x fastest), built row by
    row in column order: each row holds its -x..-z and +x..+z neighbours
    (those inside a Dirichlet grid, all of them on a periodic one) and the
    diagonal, columns sorted."""
    periodic = bc == "periodic"
    if periodic and min(shape) < 3:
        raise ValueError("periodic grids need at least 3 points per side")
    dims = len(shape)
    n = int(np.prod(shape))
    rev = tuple(reversed(shape))  # array axes (z, y, x): C order is x fastest
    idx = np.arange(n, dtype=np.int64).reshape(rev)
    strides = [int(np.prod(shape[:d])) for d in range(dims)]
    w = 2 * dims + 1
    cols = np.empty(rev + (w,), dtype=np.int64)
    keep = np.ones(rev + (w,), dtype=bool)
    vals = np.full(rev + (w,), -1.0)
    diag = np.full(n, 2.0 * dims)
    if potential is not None:
        diag = diag + np.asarray(potential, dtype=np.float64)
    for d in range(dims):
        L = shape[d]
        view = [1] * dims
        view[dims - 1 - d] = L
        c = np.arange(L).reshape(view)  # coordinate d, broadcast over the grid
        lo, hi = dims - 1 - d, dims + 1 + d  # slots of the -d and +d neighbours
        cols[..., lo] = idx - strides[d]
        cols[..., hi] = idx + strides[d]
        if periodic:
            cols[..., lo] += np.where(c == 0, L * strides[d], 0)
            cols[..., hi] -= np.where(c == L - 1, L * strides[d], 0)
        else:
            keep[..., lo] = c > 0
            keep[..., hi] = c < L - 1
    cols[..., dims] = idx
    vals[..., dims] = diag.reshape(rev)
    cols = cols.reshape(n, w)
    vals = vals.reshape(n, w)
    keep = keep.reshape(n, w)
    if periodic:  # wrapped neighbours break the slot order
        order = np.argsort(cols, axis=1, kind="stable")
        cols = np.take_along_axis(cols, order, axis=1)
        vals = np.take_along_axis(vals, order, axis=1)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.count_nonzero(keep, axis=1), out=rowptr[1:])
    A = Csr(n, rowptr, cols[keep], vals[keep])
    A.grid = tuple(shape)  # locality hint for the device kernels (row order x fastest)
    return A


def laplacian_2d(nx, ny=None, bc="dirichlet", potential=None) -> Csr:
    ny = nx if ny is None else ny
    return _laplacian((nx, ny), bc, potential)


def laplacian_3d(nx, ny=None, nz=None, bc="periodic", potential=None) -> Csr:
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _laplacian((nx, ny, nz), bc, potential)


def anderson_potential(n, W, seed) -> np.ndarray:
    """Onsite U[-W/2, W/2] in row order from stream `seed`."""
    return uniform_draws(seed, 0, n, -W / 2.0, W / 2.0)


def laplacian_2d_spectrum(nx, ny=None, bc="dirichlet") -> np.ndarray:
    ny = nx if ny is None else ny
    if bc == "dirichlet":
        a = 4.0 * np.sin(np.pi * np.arange(1, nx + 1) / (2 * (nx + 1))) ** 2
        b = 4.0 * np.sin(np.pi * np.arange(1, ny + 1) / (2 * (ny + 1))) ** 2
    else:
        a = 4.0 * np.sin(np.pi * np.arange(nx) / nx) ** 2
        b = 4.0 * np.sin(np.pi * np.arange(ny) / ny) ** 2
    return np.sort((a[:, None] + b[None, :]).ravel())


def laplacian_3d_spectrum(n, bc="periodic") -> np.ndarray:
    if bc == "periodic":
        a = 4.0 * np.sin(np.pi * np.arange(n) / n) ** 2
    else:
        a = 4.0 * np.sin(np.pi * np.arange(1, n + 1) / (2 * (n + 1))) ** 2
    return np.sort((a[:, None, None] + a[None, :, None] + a[None, None, :]).ravel())


def hofstadter_2d(L, phi=1.0 / 64, W=0.0, seed=2) -> Csr:
    """Periodic L x L lattice, hopping -1 on x-bonds, -exp(2 pi i phi x) on the
    y-bond (x,y)->(x,y+1) (and its conjugate back), onsite U[-W/2, W/2]."""
    if (L * phi) % 1.0 != 0.0:
        raise ValueError("L*phi must be an integer for a periodic gauge")
    n = L * L
    idx = np.arange(n, dtype=np.int64)
    x = idx % L
    y = idx // L
    xp = (x + 1) % L + L * y
    xm = (x - 1) % L + L * y
    yp = x + L * ((y + 1) % L)
    ym = x + L * ((y - 1) % L)
    ph = np.exp(2j * np.pi * phi * x)
    onsite = anderson_potential(n, W, seed) if W > 0 else np.zeros(n)
    rows = np.concatenate([idx, idx, idx, idx, idx])
    cols = np.concatenate([idx, xp, xm, yp, ym])
    # H[i, yp(i)] couples psi(x,y+1) into row (x,y): <x,y|H|x,y+1> = -conj(t)
    # with t the hop (x,y)->(x,y+1) = -exp(2 pi i phi x); Hermitian by construction.
    vals = np.concatenate([onsite.astype(np.complex128), -np.ones(n, complex), -np.ones(n, complex),
                           -np.conj(ph), -ph])
    return _from_coo(n, rows, cols, vals)


def dense_qlq(n, lam=None, seed=3, dtype=np.float64) -> np.ndarray:
    """A = Q diag(lam) Q^T, Q = qr(G).Q diag(sign(diag R)), G ~ N(0,1) by
    Box-Muller from stream `seed`.  Returns a column-major (Fortran) array."""
    if lam is None:
        lam = np.linspace(0.0, 1.0, n) ** 2
    u = unit_draws(seed, 0, n * n + (n * n) % 2)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.empty(u.shape[0])
    g[0::2] = r * np.cos(2 * np.pi * u2)
    g[1::2] = r * np.sin(2 * np.pi * u2)
    G = g[: n * n].reshape(n, n, order="F")
    Q, R = np.linalg.qr(G)
    Q = Q * np.sign(np.diag(R))[None, :]
    A = (Q * lam[None, :]) @ Q.T
    A = 0.5 * (A + A.T)
    return np.asfortranarray(A.astype(dtype))
