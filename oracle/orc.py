"""ctypes binding of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py as the checker.  The functions restate the
reference (see rc_oracle.hpp for the file:line each one follows).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")


class Params(C.Structure):
    _fields_ = [
        ("k", C.c_long), ("inc", C.c_long), ("buf", C.c_long), ("tol", C.c_double),
        ("seed", C.c_ulonglong), ("stop", C.c_int), ("interval", C.c_int),
        ("max_inner", C.c_long), ("deg_max", C.c_long), ("rep", C.c_long),
        ("lambda_max", C.c_double), ("lanczos_itmax", C.c_long), ("maxmin_tol", C.c_double),
        ("fixed_iteration", C.c_int), ("pad", C.c_int), ("out_cap", C.c_long),
    ]


class Result(C.Structure):
    _fields_ = [
        ("found", C.c_long), ("lanczos_iters", C.c_long), ("min_est", C.c_double),
        ("max_est", C.c_double), ("guard", C.c_double), ("count_ortho", C.c_long),
        ("count_apply", C.c_long), ("count_eig", C.c_long), ("final_maxres", C.c_double),
        ("final_maxdiff", C.c_double), ("final_maxgap", C.c_double), ("trace_len", C.c_long),
        ("rng_counter", C.c_ulonglong), ("t_lanczos", C.c_double), ("t_filter", C.c_double),
        ("t_ortho", C.c_double), ("t_deflate", C.c_double), ("t_rr", C.c_double),
        ("n_ortho", C.c_long), ("n_deflate", C.c_long), ("n_rr", C.c_long),
        ("degree_sum", C.c_long),
    ]


STOP = {"DEFAULT": 0, "COMBINATION": 1, "EIGENVALUE_ONLY": 2, "RESIDUAL_ONLY": 3}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        d, l, p = C.c_double, C.c_long, C.c_void_p
        L.orc_last_error.restype = C.c_char_p
        L.orc_unit_draw.restype = d
        L.orc_unit_draw.argtypes = [C.c_ulonglong, C.c_ulonglong]
        L.orc_fill_sym.argtypes = [C.c_ulonglong, C.c_ulonglong, p, l]
        for name in ("orc_cpoly_pm", "orc_cpoly_pm_prime"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [l, d]
        for name in ("orc_cpoly_xstar", "orc_cpoly_upper_xstar", "orc_cpoly_pm_prime_first_root"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [l]
        L.orc_cpoly_inverse_pm.restype = d
        L.orc_cpoly_inverse_pm.argtypes = [l, d]
        L.orc_cpoly_star_degree.argtypes = [d, d, d, l, C.POINTER(l), C.POINTER(d)]
        L.orc_cpoly_eval_polynomial.restype = d
        L.orc_cpoly_eval_polynomial.argtypes = [l, l, d, d, d]
        L.orc_tri_eig.restype = d
        L.orc_tri_eig.argtypes = [p, p, l, l, d]
        L.orc_jacobi_f64.argtypes = [p, l, p, p]
        L.orc_jacobi_c128.argtypes = [p, l, p, p]
        L.orc_mgs_f64.argtypes = [p, l, l]
        L.orc_filter_csr_f64.argtypes = [l, p, p, p, p, l, l, l, d, d]
        L.orc_filter_csr_c128.argtypes = [l, p, p, p, p, l, l, l, d, d]
        L.orc_filter_dense_f64.argtypes = [l, p, p, l, l, l, d, d]
        L.orc_lanczos_csr_f64.argtypes = [l, p, p, p, C.c_ulonglong, d, C.c_int, C.c_int, l,
                                          C.POINTER(d), C.POINTER(d), C.POINTER(l)]
        for name in ("orc_solve_csr_f64", "orc_solve_csr_c128"):
            getattr(L, name).argtypes = [l, p, p, p, C.POINTER(Params), C.POINTER(Result),
                                         p, p, p, p, l]
        L.orc_solve_dense_f64.argtypes = [l, p, C.POINTER(Params), C.POINTER(Result), p, p, p, p, l]
        L.orc_sizeof_params.restype = l
        L.orc_sizeof_result.restype = l
        assert L.orc_sizeof_params() == C.sizeof(Params)
        assert L.orc_sizeof_result() == C.sizeof(Result)
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        msg = lib().orc_last_error().decode()
        if rc == 2:
            raise ValueError(msg)
        raise RuntimeError(msg)


def sym_draws(seed, start, count):
    out = np.empty(count)
    lib().orc_fill_sym(seed, start, _ptr(out), count)
    return out


def cpoly_pm(m, x):
    return lib().orc_cpoly_pm(m, x)


def cpoly_pm_prime(m, x):
    return lib().orc_cpoly_pm_prime(m, x)


def cpoly_xstar(m):
    return lib().orc_cpoly_xstar(m)


def cpoly_upper_xstar(m):
    return lib().orc_cpoly_upper_xstar(m)


def cpoly_inverse_pm(m, y):
    return lib().orc_cpoly_inverse_pm(m, y)


def cpoly_star_degree(shift, lambda_bound, lambda_star, deg_max=200):
    deg = C.c_long(0)
    bound = C.c_double(0.0)
    lib().orc_cpoly_star_degree(shift, lambda_bound, lambda_star, deg_max, C.byref(deg), C.byref(bound))
    return deg.value, bound.value


def cpoly_eval_polynomial(m, rep, lmax, shift, lam):
    return lib().orc_cpoly_eval_polynomial(m, rep, lmax, shift, lam)


def tri_eig(D, E, idx, abstol=1e-14):
    D = np.ascontiguousarray(D, dtype=np.float64)
    E = np.ascontiguousarray(E, dtype=np.float64)
    return lib().orc_tri_eig(_ptr(D), _ptr(E), D.shape[0], idx, abstol)


def jacobi(H):
    """Symmetric (float64) or Hermitian (complex128) eigensystem, ascending."""
    H = np.asfortranarray(H)
    n = H.shape[0]
    w = np.empty(n)
    if np.iscomplexobj(H):
        H = H.astype(np.complex128, order="F")
        V = np.empty((n, n), dtype=np.complex128, order="F")
        lib().orc_jacobi_c128(_ptr(H), n, _ptr(w), _ptr(V))
    else:
        H = H.astype(np.float64, order="F")
        V = np.empty((n, n), order="F")
        lib().orc_jacobi_f64(_ptr(H), n, _ptr(w), _ptr(V))
    return w, V


def mgs(X):
    X = np.array(X, dtype=np.float64, order="F")
    lib().orc_mgs_f64(_ptr(X), X.shape[0], X.shape[1])
    return X


def filter_csr(A, X, m, lambda_max, shift, rep=1):
    """P_m(A')^rep X for a column-major block X (n x b)."""
    if A.is_complex:
        Xc = np.array(X, dtype=np.complex128, order="F")
        _check(lib().orc_filter_csr_c128(A.n, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val), _ptr(Xc),
                                         Xc.shape[1], m, rep, lambda_max, shift))
        return Xc
    Xf = np.array(X, dtype=np.float64, order="F")
    _check(lib().orc_filter_csr_f64(A.n, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val), _ptr(Xf),
                                    Xf.shape[1], m, rep, lambda_max, shift))
    return Xf


def filter_dense(A, X, m, lambda_max, shift, rep=1):
    A = np.asfortranarray(A, dtype=np.float64)
    Xf = np.array(X, dtype=np.float64, order="F")
    _check(lib().orc_filter_dense_f64(A.shape[0], _ptr(A), _ptr(Xf), Xf.shape[1], m, rep,
                                      lambda_max, shift))
    return Xf


def lanczos_bounds(A, seed=0, tol=1e-6, max_flag=True, min_flag=False, itmax=10000):
    mn, mx, it = C.c_double(), C.c_double(), C.c_long()
    _check(lib().orc_lanczos_csr_f64(A.n, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val), seed, tol,
                                     int(max_flag), int(min_flag), itmax, C.byref(mn), C.byref(mx),
                                     C.byref(it)))
    return mn.value, mx.value, it.value


def solve(A, k, tol, inc, buf, seed=0, stop="COMBINATION", interval=False, lambda_max=0.0,
          max_inner=0, deg_max=0, rep=0, lanczos_itmax=0, maxmin_tol=0.0, fixed_iteration=False,
          want_vectors=False, trace_cap=100000):
    """RayleighChebyshevLM::getMinEigenSystem (or getMinIntervalEigenSystem)."""
    dense = isinstance(A, np.ndarray)
    n = A.shape[0] if dense else A.n
    cap = max(k, inc + buf, 1)
    cplx = (not dense) and A.is_complex
    p = Params(k=k, inc=inc, buf=buf, tol=tol, seed=seed, stop=STOP[stop.upper()],
               interval=int(interval), max_inner=max_inner, deg_max=deg_max, rep=rep,
               lambda_max=lambda_max, lanczos_itmax=lanczos_itmax, maxmin_tol=maxmin_tol,
               fixed_iteration=int(fixed_iteration), pad=0, out_cap=cap)
    r = Result()
    evals = np.zeros(cap)
    resid = np.zeros(cap)
    evecs = np.zeros((n, cap), dtype=np.complex128 if cplx else np.float64, order="F") if want_vectors else None
    trace = np.zeros((trace_cap, 10))
    ev_ptr = _ptr(evecs) if want_vectors else None
    if dense:
        Af = np.asfortranarray(A, dtype=np.float64)
        rc = lib().orc_solve_dense_f64(n, _ptr(Af), C.byref(p), C.byref(r), _ptr(evals), ev_ptr,
                                       _ptr(resid), _ptr(trace), trace_cap)
    elif cplx:
        rc = lib().orc_solve_csr_c128(n, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val), C.byref(p),
                                      C.byref(r), _ptr(evals), ev_ptr, _ptr(resid), _ptr(trace),
                                      trace_cap)
    else:
        rc = lib().orc_solve_csr_f64(n, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val), C.byref(p),
                                     C.byref(r), _ptr(evals), ev_ptr, _ptr(resid), _ptr(trace),
                                     trace_cap)
    _check(rc)
    f = r.found
    out = {
        "found": f,
        "eigenvalues": evals[:f].copy(),
        "residuals": resid[:f].copy(),
        "lanczos_iters": r.lanczos_iters,
        "min_est": r.min_est,
        "max_est": r.max_est,
        "guard": r.guard,
        "count": {"ortho": r.count_ortho, "OpApply": r.count_apply, "eigenvalue": r.count_eig},
        "final": {"maxResidual": r.final_maxres, "maxEigValueDifference": r.final_maxdiff,
                  "maxRelEigValueGap": r.final_maxgap},
        "trace": trace[: min(r.trace_len, trace_cap)].copy(),
        "rng_counter": r.rng_counter,
        "timing": {"lanczos": r.t_lanczos, "filter": r.t_filter, "ortho": r.t_ortho,
                   "deflate": r.t_deflate, "rr": r.t_rr, "n_ortho": r.n_ortho,
                   "n_deflate": r.n_deflate, "n_rr": r.n_rr, "degree_sum": r.degree_sum},
    }
    if want_vectors:
        out["eigenvectors"] = evecs[:, :f].copy()
    return out


def cpoly_pm_prime_first_root(m):
    return lib().orc_cpoly_pm_prime_first_root(m)
