"""Operator ingest on the host: Matrix Market files and 64-bit CSR indices
(librcg.so, no device needed for the readers), checked against scipy."""
import ctypes as C

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from rayleighchebyshevgpgpu_amd import capi
from rayleighchebyshevgpgpu_amd import operators as ops


def read_mm(path):
    L = capi.lib()
    n, nnz, dt = C.c_int64(), C.c_int64(), C.c_int()
    capi.check(L.rcg_matrix_market_info(str(path).encode(), C.byref(n), C.byref(nnz), C.byref(dt)))
    rp = np.empty(n.value + 1, dtype=np.int32)
    col = np.empty(nnz.value, dtype=np.int32)
    val = np.empty(nnz.value, dtype=np.complex128 if dt.value == capi.C128 else np.float64)
    capi.check(L.rcg_matrix_market_csr(str(path).encode(), rp.ctypes.data, col.ctypes.data,
                                       val.ctypes.data))
    return sp.csr_matrix((val, col, rp), shape=(n.value, n.value)), dt.value


def test_symmetric_real(tmp_path):
    A = ops.laplacian_2d(12, potential=np.linspace(0, 1, 144)).to_scipy()
    path = tmp_path / "a.mtx"
    scipy.io.mmwrite(str(path), sp.tril(A).tocoo(), symmetry="symmetric")
    B, dt = read_mm(path)
    assert dt == capi.F64
    assert (abs(A - B)).max() < 1e-15 and B.has_sorted_indices


def test_hermitian_complex(tmp_path):
    A = ops.hofstadter_2d(16, phi=1 / 8, W=1.0).to_scipy()
    path = tmp_path / "h.mtx"
    scipy.io.mmwrite(str(path), sp.tril(A).tocoo(), symmetry="hermitian")
    B, dt = read_mm(path)
    assert dt == capi.C128
    assert (abs(A - B)).max() < 1e-15


def test_general_pattern_and_duplicates(tmp_path):
    path = tmp_path / "p.mtx"
    path.write_text("%%MatrixMarket matrix coordinate pattern general\n% comment\n3 3 4\n"
                    "1 1\n3 2\n1 1\n2 3\n")
    B, dt = read_mm(path)
    assert B.toarray().tolist() == [[2, 0, 0], [0, 0, 1], [0, 1, 0]]


def test_bad_files_report_errors(tmp_path):
    L = capi.lib()
    n, nnz, dt = C.c_int64(), C.c_int64(), C.c_int()
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    assert L.rcg_matrix_market_info(str(bad).encode(), C.byref(n), C.byref(nnz), C.byref(dt)) != 0
    assert "coordinate" in capi.last_error()
    rect = tmp_path / "rect.mtx"
    rect.write_text("%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1.0\n")
    assert L.rcg_matrix_market_info(str(rect).encode(), C.byref(n), C.byref(nnz), C.byref(dt)) != 0
    assert "square" in capi.last_error()
    oob = tmp_path / "oob.mtx"
    oob.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    assert L.rcg_matrix_market_info(str(oob).encode(), C.byref(n), C.byref(nnz), C.byref(dt)) != 0
    missing = tmp_path / "none.mtx"
    assert L.rcg_matrix_market_info(str(missing).encode(), C.byref(n), C.byref(nnz), C.byref(dt)) != 0
    assert "cannot open" in capi.last_error()


def test_csr64_rejects_bad_indices():
    # validation happens on the host before any device work; a NULL context is rejected first
    L = capi.lib()
    rp = np.array([0, 1, 5], dtype=np.int64)
    col = np.array([0, 1], dtype=np.int64)
    val = np.ones(2)
    out = C.c_void_p()
    assert L.rcg_operator_csr64_create(None, 2, 2, rp.ctypes.data, col.ctypes.data, val.ctypes.data,
                                       capi.F64, C.byref(out)) != 0
