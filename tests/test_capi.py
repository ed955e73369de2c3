"""The C-ABI library loads and exports every symbol include/rcg_c_api.h
declares; host-side errors surface as status codes (no GPU needed)."""
import ctypes as C

import pytest

from rayleighchebyshevgpgpu_amd import capi


def test_library_exports_every_declared_function():
    L = capi.lib()
    names = capi.declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def test_struct_layouts_match():
    L = capi.lib()
    assert L.rcg_sizeof_solve_params() == C.sizeof(capi.SolveParams)
    assert L.rcg_sizeof_solve_result() == C.sizeof(capi.SolveResult)


def test_version_names_target():
    assert b"gfx950" in capi.lib().rcg_version()


def test_null_arguments_return_status_not_exit():
    L = capi.lib()
    assert L.rcg_context_create(0, None, None) == capi.RCG_OK + 1
    assert "NULL" in capi.last_error() or "null" in capi.last_error().lower()
    assert L.rcg_block_create(None, 10, 2, capi.F64, None) == 1
    assert L.rcg_solve_lowest(None, None, None, None, None, 0, None, None, 0) == 1


def test_small_eigensystem_host_path():
    """rcg_small_eigensystem runs on the host (no device needed)."""
    import numpy as np
    rng = np.random.default_rng(0)
    # Jacobi below 48, Householder + QL from 48 (with a clustered spectrum too)
    cases = [(1, False), (2, False), (5, False), (33, False), (48, False), (100, False), (64, True)]
    for n, clustered in cases:
        G = rng.standard_normal((n, n))
        if clustered:
            Q, _ = np.linalg.qr(G)
            G = (Q * np.repeat([1.0, 2.0, 2.0 + 1e-9, 5.0], n // 4)) @ Q.T
        H = np.asfortranarray((G + G.T) / 2)
        w = np.empty(n)
        Z = np.empty((n, n), order="F")
        capi.check(capi.lib().rcg_small_eigensystem(n, H.ctypes.data_as(C.c_void_p),
                                                    w.ctypes.data_as(C.c_void_p),
                                                    Z.ctypes.data_as(C.c_void_p)))
        assert np.allclose(w, np.linalg.eigvalsh(H), atol=1e-13 * max(1.0, n / 10))
        assert np.max(np.abs(H @ Z - Z * w)) < 1e-12 * max(1.0, n / 10)
        assert np.max(np.abs(Z.T @ Z - np.eye(n))) < 1e-13


def test_python_mirror_rejects_bad_stop_string():
    from rayleighchebyshevgpgpu_amd.solver import RayleighChebyshevLM
    rc = RayleighChebyshevLM()
    rc.setStopCondition("residual_only")
    assert rc.stopCondition == capi.STOP["RESIDUAL_ONLY"]
    rc.setStopCondition("DEFAULT")
    assert rc.stopCondition == capi.STOP["COMBINATION"]
    with pytest.raises(RuntimeError):
        rc.setStopCondition("FASTEST")
