"""ctypes binding of librcg.so (include/rcg_c_api.h).

Loading fails loudly when the library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librcg.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "rcg_c_api.h")

RCG_OK = 0
F64, F32, C128, C64 = 0, 1, 2, 3
STOP = {"DEFAULT": 0, "COMBINATION": 1, "EIGENVALUE_ONLY": 2, "RESIDUAL_ONLY": 3}
MEM_HOST, MEM_DEVICE = 0, 1


class SolveParams(C.Structure):
    _fields_ = [
        ("k", C.c_int64), ("inc", C.c_int64), ("buf", C.c_int64), ("tol", C.c_double),
        ("seed", C.c_uint64), ("stop", C.c_int32), ("interval", C.c_int32),
        ("max_inner", C.c_int64), ("deg_max", C.c_int64), ("rep", C.c_int64),
        ("lambda_max", C.c_double), ("lanczos_itmax", C.c_int64), ("maxmin_tol", C.c_double),
        ("fixed_iteration", C.c_int32), ("filter_dtype", C.c_int32), ("refine", C.c_int32),
        ("verbose", C.c_int32), ("use_bounds", C.c_int32), ("nonrandom_start", C.c_int32),
        ("min_eig", C.c_double), ("max_eig", C.c_double),
    ]


class SolveResult(C.Structure):
    _fields_ = [
        ("found", C.c_int64), ("lanczos_iters", C.c_int64), ("min_est", C.c_double),
        ("max_est", C.c_double), ("guard", C.c_double), ("count_ortho", C.c_int64),
        ("count_apply", C.c_int64), ("count_eig", C.c_int64), ("final_maxres", C.c_double),
        ("final_maxdiff", C.c_double), ("final_maxgap", C.c_double), ("trace_len", C.c_int64),
        ("rng_counter", C.c_uint64), ("filter_degree_sum", C.c_int64), ("t_total", C.c_double),
        ("t_lanczos", C.c_double), ("t_filter", C.c_double), ("t_ortho", C.c_double),
        ("t_rr", C.c_double), ("t_filter_device", C.c_double), ("filter_bytes", C.c_double),
        ("filter_panel", C.c_int64), ("filter_launches", C.c_int64), ("t_eig", C.c_double),
    ]


# int apply(void* user, const void* in, void* out, int64_t n, int64_t cols,
#           int64_t ld, double alpha, double beta, void* hip_stream)
APPLY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64,
                       C.c_int64, C.c_double, C.c_double, C.c_void_p)


class RcgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rcg status {code}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"librcg.so is not built ({LIB_PATH}); run "
                           "python -m rayleighchebyshevgpgpu_amd.build")
    L = C.CDLL(LIB_PATH)
    vp, i64, u64, d, i32 = C.c_void_p, C.c_int64, C.c_uint64, C.c_double, C.c_int
    P = C.POINTER
    sig = {
        "rcg_get_last_error": ([C.c_char_p, C.c_size_t], i32),
        "rcg_version": ([], C.c_char_p),
        "rcg_device_count": ([P(i32)], i32),
        "rcg_context_create": ([i32, vp, P(vp)], i32),
        "rcg_context_destroy": ([vp], i32),
        "rcg_context_synchronize": ([vp], i32),
        "rcg_memory_trim": ([], i32),
        "rcg_block_create": ([vp, i64, i64, i32, P(vp)], i32),
        "rcg_block_destroy": ([vp], i32),
        "rcg_block_upload": ([vp, vp, i64, i64], i32),
        "rcg_block_download": ([vp, vp, i64, i64], i32),
        "rcg_block_copy": ([vp, vp], i32),
        "rcg_block_randomize": ([vp, i64, i64, u64, P(u64)], i32),
        "rcg_block_info": ([vp, P(i64), P(i64), P(i64), P(i32), P(vp)], i32),
        "rcg_operator_csr_create": ([vp, i64, i64, vp, vp, vp, i32, i32, P(vp)], i32),
        "rcg_operator_destroy": ([vp], i32),
        "rcg_operator_dense_create": ([vp, i64, vp, i64, i32, P(vp)], i32),
        "rcg_operator_csr64_create": ([vp, i64, i64, vp, vp, vp, i32, P(vp)], i32),
        "rcg_matrix_market_info": ([C.c_char_p, P(i64), P(i64), P(i32)], i32),
        "rcg_matrix_market_csr": ([C.c_char_p, vp, vp, vp], i32),
        "rcg_operator_from_matrix_market": ([vp, C.c_char_p, P(vp)], i32),
        "rcg_operator_set_grid": ([vp, i64, i64, i64], i32),
        "rcg_operator_csr_share": ([vp, vp, i32, P(vp)], i32),
        "rcg_operator_callback_create": ([vp, i64, i32, APPLY_FN, vp, P(vp)], i32),
        "rcg_operator_apply": ([vp, vp, vp, d, d, i64], i32),
        "rcg_cheb_filter": ([vp, vp, vp, i64, i64, i64, d, d], i32),
        "rcg_cheb_filter_mixed": ([vp, vp, vp, i64, i64, i64, d, d, vp, vp], i32),
        "rcg_cheb_filter_batch": ([i64, vp, vp, vp, vp, vp, vp, vp], i32),
        "rcg_orthonormalize": ([vp, vp, i64], i32),
        "rcg_deflate": ([vp, i64, vp, i64], i32),
        "rcg_project": ([vp, vp, vp, i64, vp], i32),
        "rcg_rotate": ([vp, vp, i64, vp], i32),
        "rcg_small_eigensystem": ([i64, vp, vp, vp], i32),
        "rcg_residuals": ([vp, vp, vp, i64, vp, vp], i32),
        "rcg_residuals_noalloc": ([vp, vp, vp, i64, vp, vp], i32),
        "rcg_lanczos_bounds": ([vp, d, i32, i64, u64, P(u64), P(d), P(d), P(i64)], i32),
        "rcg_solve_lowest": ([vp, P(SolveParams), P(SolveResult), vp, vp, i64, vp, vp, i64], i32),
        "rcg_server_create": ([vp, i64, i32, i32, P(vp)], i32),
        "rcg_server_submit": ([vp, vp, P(SolveParams), P(i64)], i32),
        "rcg_server_wait": ([vp, i64, P(SolveResult), vp, vp, i64, vp], i32),
        "rcg_server_vectors": ([vp, i64, P(vp)], i32),
        "rcg_server_release": ([vp, i64], i32),
        "rcg_server_stats": ([vp, vp], i32),
        "rcg_server_submit_warm": ([vp, vp, vp, i64, vp], i32),
        "rcg_server_filter_profile": ([vp, vp], i32),
        "rcg_server_destroy": ([vp], i32),
        "rcg_sizeof_solve_params": ([], i64),
        "rcg_sizeof_solve_result": ([], i64),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    if L.rcg_sizeof_solve_params() != C.sizeof(SolveParams):
        raise RuntimeError("rcg_solve_params layout mismatch")
    if L.rcg_sizeof_solve_result() != C.sizeof(SolveResult):
        raise RuntimeError("rcg_solve_result layout mismatch")
    _lib = L
    return L


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    lib().rcg_get_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(code: int) -> None:
    if code != RCG_OK:
        raise RcgError(code, last_error())


def declared_functions() -> list[str]:
    """Names of the functions include/rcg_c_api.h declares."""
    import re
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rcg_[a-z0-9_]+)\s*\(", text)))
