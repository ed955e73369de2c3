"""The synthetic stencil operators: the row-by-row CSR builder gives the same
arrays as assembling the bonds as COO triplets and converting with scipy
(sorted columns, no duplicates), for Dirichlet and periodic grids of one to
three dimensions, including the smallest periodic side (3) and uneven sides."""
import numpy as np
import pytest
import scipy.sparse as sp

from rayleighchebyshevgpgpu_amd import operators as ops


def coo_reference(shape, periodic, pot):
    dims = len(shape)
    n = int(np.prod(shape))
    idx = np.arange(n)
    rows, cols = [idx], [idx]
    vals = [np.full(n, 2.0 * dims) + pot]
    rem = idx.copy()
    stride = 1
    for d in range(dims):
        L = shape[d]
        c = rem % L
        rem = rem // L
        for step in (-1, 1):
            t = c + step
            ok = np.ones(n, bool) if periodic else (t >= 0) & (t < L)
            t = t % L
            rows.append(idx[ok])
            cols.append((idx + (t - c) * stride)[ok])
            vals.append(-np.ones(int(ok.sum())))
        stride *= L
    m = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n)).tocsr()
    m.sort_indices()
    return m


@pytest.mark.parametrize("shape", [(7,), (3, 3), (4, 9), (3, 3, 3), (5, 4, 6), (16, 16, 16)])
@pytest.mark.parametrize("bc", ["dirichlet", "periodic"])
def test_stencil_csr_matches_coo_assembly(shape, bc):
    n = int(np.prod(shape))
    pot = ops.anderson_potential(n, 4.0, 5)
    A = ops._laplacian(shape, bc, pot)
    ref = coo_reference(shape, bc == "periodic", pot)
    assert np.array_equal(A.rowptr, ref.indptr)
    assert np.array_equal(A.col, ref.indices)
    assert np.array_equal(A.val, ref.data)
    assert A.grid == tuple(shape)
