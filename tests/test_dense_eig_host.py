"""Host Hermitian eigensolver of the Rayleigh-Ritz step (rcg_dense.hpp) on
complex matrices: cyclic Jacobi below 48, Householder on the lower triangle +
implicit QL from 48.  Pins: numpy.linalg.eigh eigenvalues; the residual
||H Z - Z diag(w)|| and the unitarity of Z; a clustered spectrum.  The
harness tests/cpp/dense_eig.cpp is compiled with g++ for the test."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("eig") / "dense_eig")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fcx-limited-range",
                    os.path.join(HERE, "cpp", "dense_eig.cpp"), "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("n,clustered", [(20, False), (48, False), (96, False), (64, True)])
def test_hermitian_eigensystem(harness, tmp_path, n, clustered):
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    if clustered:
        Q, _ = np.linalg.qr(G)
        G = (Q * np.repeat([1.0, 2.0, 2.0 + 1e-9, 5.0], n // 4)) @ Q.conj().T
    H = (G + G.conj().T) / 2
    fin, fout = tmp_path / "h.bin", tmp_path / "z.bin"
    with open(fin, "wb") as f:
        f.write(np.int64(n).tobytes())
        f.write(np.asfortranarray(H).tobytes(order="F"))
    subprocess.run([harness, str(fin), str(fout)], check=True)
    raw = np.fromfile(fout, dtype=np.float64)
    w = raw[:n]
    Z = raw[n:].view(np.complex128).reshape((n, n), order="F")
    scale = max(1.0, n / 10)
    assert np.allclose(w, np.linalg.eigvalsh(H), atol=1e-13 * scale)
    assert np.max(np.abs(H @ Z - Z * w)) < 1e-12 * scale
    assert np.max(np.abs(Z.conj().T @ Z - np.eye(n))) < 1e-13 * scale
