"""Python host side over the C ABI, mirroring the reference's interface.

The reference is a C++ template library used as

    RayleighChebyshevLM<Vtype, Otype, RandOp> rc;
    rc.setStopCondition("RESIDUAL_ONLY");
    long found = rc.getMinEigenSystem(eigCount, tol, inc, buf, vStart, oP, randOp,
                                      eigValues, eigVectors);
    rc.getEigVectorResiduals(); rc.getCountData(); rc.getFinalData();

(RayleighChebyshevLM.h:281-472, 693-806).  ``RayleighChebyshevLM`` below has
the same method names, argument meaning and error behaviour, over device
blocks (``Block``, the AmatrixClassLMCuda role), a device CSR operator
(``CsrOperator``, the user Otype) and ``RandomizeOp`` (ArandomizeOpClassLMCuda,
with a seed).  Everything runs on the GPU through librcg.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, lib


class Context:
    """Device + stream (rcg_context)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        h = C.c_void_p()
        check(lib().rcg_context_create(device, C.c_void_p(stream) if stream else None, C.byref(h)))
        self.h = h
        self.device = device

    def synchronize(self):
        check(lib().rcg_context_synchronize(self.h))

    def close(self):
        if self.h:
            lib().rcg_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Block:
    """N x cols device block (AmatrixClassLMCuda role); host side column-major."""

    def __init__(self, n: int, cols: int, dtype=np.float64, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.n, self.cols = int(n), int(cols)
        self.dtype = np.dtype(dtype)
        codes = {np.dtype(np.float64): capi.F64, np.dtype(np.float32): capi.F32,
                 np.dtype(np.complex128): capi.C128}
        if self.dtype not in codes:
            raise TypeError(f"blocks are float64, float32 or complex128, not {self.dtype}")
        code = codes[self.dtype]
        h = C.c_void_p()
        check(lib().rcg_block_create(self.ctx.h, self.n, self.cols, code, C.byref(h)))
        self.h = h
        ld = C.c_int64()
        check(lib().rcg_block_info(self.h, None, None, C.byref(ld), None, None))
        self.ld = ld.value

    @classmethod
    def from_host(cls, X: np.ndarray, ctx: Context | None = None) -> "Block":
        X = np.asfortranarray(X)
        b = cls(X.shape[0], X.shape[1], X.dtype, ctx)
        b.upload(X)
        return b

    def upload(self, X: np.ndarray):
        X = np.asfortranarray(X, dtype=self.dtype)
        check(lib().rcg_block_upload(self.h, X.ctypes.data_as(C.c_void_p), X.shape[0], X.shape[1]))

    def download(self, cols: int | None = None) -> np.ndarray:
        cols = self.cols if cols is None else cols
        X = np.empty((self.n, cols), dtype=self.dtype, order="F")
        check(lib().rcg_block_download(self.h, X.ctypes.data_as(C.c_void_p), self.n, cols))
        return X

    def randomize(self, col0: int, ncols: int, seed: int, counter: int) -> int:
        c = C.c_uint64(counter)
        check(lib().rcg_block_randomize(self.h, col0, ncols, seed, C.byref(c)))
        return c.value

    def close(self):
        if getattr(self, "h", None):
            lib().rcg_block_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CsrOperator:
    """Device CSR operator (the reference's user Otype with
    apply(in, out, alpha, beta), LanczosCpolyOperatorLMCuda.h:236,245)."""

    def __init__(self, csr, ctx: Context | None = None, grid=None):
        self.ctx = ctx or default_context()
        self.n, self.nnz = csr.n, csr.nnz
        self.is_complex = bool(csr.is_complex)
        self.dtype = np.dtype(np.complex128 if self.is_complex else np.float64)
        val = np.ascontiguousarray(csr.val, dtype=self.dtype)
        rowptr = np.ascontiguousarray(csr.rowptr, dtype=np.int32)
        col = np.ascontiguousarray(csr.col, dtype=np.int32)
        h = C.c_void_p()
        check(lib().rcg_operator_csr_create(
            self.ctx.h, csr.n, csr.nnz, rowptr.ctypes.data_as(C.c_void_p),
            col.ctypes.data_as(C.c_void_p), val.ctypes.data_as(C.c_void_p),
            capi.C128 if self.is_complex else capi.F64, capi.MEM_HOST, C.byref(h)))
        self.h = h
        self._keep = None
        grid = grid if grid is not None else getattr(csr, "grid", None)
        if grid is not None:
            nx, ny, nz = (tuple(grid) + (1, 1))[:3]
            check(lib().rcg_operator_set_grid(self.h, nx, ny, nz))

    def apply(self, X: Block, Y: Block, alpha: float = 1.0, beta: float = 0.0, cols=None):
        check(lib().rcg_operator_apply(self.h, X.h, Y.h, alpha, beta, X.cols if cols is None else cols))

    @classmethod
    def from_matrix_market(cls, path: str, ctx: Context | None = None) -> "CsrOperator":
        """An operator read from a Matrix Market coordinate file by the library."""
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        n, nnz, dt = C.c_int64(), C.c_int64(), C.c_int()
        check(lib().rcg_matrix_market_info(str(path).encode(), C.byref(n), C.byref(nnz), C.byref(dt)))
        self.n, self.nnz = n.value, nnz.value
        self.is_complex = dt.value == capi.C128
        self.dtype = np.dtype(np.complex128 if self.is_complex else np.float64)
        h = C.c_void_p()
        check(lib().rcg_operator_from_matrix_market(self.ctx.h, str(path).encode(), C.byref(h)))
        self.h = h
        return self

    def share(self, val) -> "CsrOperator":
        """An operator with this operator's sparsity pattern (row pointers and
        columns shared on the device) and its own values (nnz entries)."""
        o = CsrOperator.__new__(CsrOperator)
        o.ctx, o.n, o.nnz = self.ctx, self.n, self.nnz
        o.is_complex, o.dtype = self.is_complex, self.dtype
        val = np.ascontiguousarray(val, dtype=self.dtype)
        if val.shape != (self.nnz,):
            raise ValueError("one value per stored entry of the pattern")
        h = C.c_void_p()
        check(lib().rcg_operator_csr_share(self.h, val.ctypes.data_as(C.c_void_p), capi.MEM_HOST,
                                           C.byref(h)))
        o.h = h
        o._keep = self  # the pattern's owner outlives the sharer
        return o

    @classmethod
    def from_csr64(cls, n, rowptr, col, val, ctx: Context | None = None) -> "CsrOperator":
        """An operator from 64-bit CSR index arrays (narrowed by the library)."""
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int64)
        self.is_complex = np.iscomplexobj(val)
        self.dtype = np.dtype(np.complex128 if self.is_complex else np.float64)
        val = np.ascontiguousarray(val, dtype=self.dtype)
        self.n, self.nnz = int(n), int(col.shape[0])
        h = C.c_void_p()
        check(lib().rcg_operator_csr64_create(self.ctx.h, self.n, self.nnz, rowptr.ctypes.data,
                                              col.ctypes.data, val.ctypes.data,
                                              capi.C128 if self.is_complex else capi.F64,
                                              C.byref(h)))
        self.h = h
        return self

    def close(self):
        if getattr(self, "h", None):
            lib().rcg_operator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def check_op(code: int, op) -> None:
    """check() that re-raises a caller operator's own exception as the cause."""
    err = getattr(op, "error", None)
    if code != capi.RCG_OK and err is not None:
        op.error = None
        raise capi.RcgError(code, capi.last_error()) from err
    check(code)


class CallbackOperator:
    """A caller's operator as code (the reference's Otype::apply(in, out,
    alpha, beta), LanczosCpolyOperatorLMCuda.h:236,245).

    ``apply(in_ptr, out_ptr, n, cols, ld, alpha, beta, stream)`` gets device
    pointers (ints) to column-major n x cols blocks (leading dimension ld,
    float64 or complex128) and the library stream (int); it must make
    out = alpha*A*in + beta*out on that stream, typically by calling compiled
    code of its own.  ``CallbackOperator.from_c`` takes such a function already
    compiled (an ``rcg_apply_fn`` pointer and its user pointer) and keeps
    Python out of the solve.  The operator must be Hermitian.  An exception in
    ``apply`` fails the library call and is re-raised as its cause."""

    def __init__(self, n: int, apply=None, dtype=np.float64, ctx: Context | None = None,
                 c_function=None, c_user=None):
        self.ctx = ctx or default_context()
        self.n = int(n)
        self.nnz = 0
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float64, np.complex128):
            raise ValueError("a caller's operator is float64 or complex128")
        self.is_complex = self.dtype == np.complex128
        self.error = None
        if c_function is not None:
            self._cb = capi.APPLY_FN(c_function) if isinstance(c_function, int) else c_function
            user = c_user
        else:
            def trampoline(user, pin, pout, n, cols, ld, alpha, beta, stream):
                try:
                    apply(pin, pout, n, cols, ld, alpha, beta, stream)
                    return 0
                except BaseException as e:  # reported through the library's status
                    self.error = e
                    return 1
            self._cb = capi.APPLY_FN(trampoline)  # kept alive with the operator
            user = None
        h = C.c_void_p()
        check(lib().rcg_operator_callback_create(self.ctx.h, self.n,
                                                 capi.C128 if self.is_complex else capi.F64,
                                                 self._cb, user, C.byref(h)))
        self.h = h

    @classmethod
    def from_c(cls, n: int, function, user, dtype=np.float64, ctx: Context | None = None):
        """A compiled apply: ``function`` an rcg_apply_fn (address or ctypes
        function pointer), ``user`` its first argument."""
        return cls(n, dtype=dtype, ctx=ctx, c_function=function, c_user=user)

    def apply(self, X: "Block", Y: "Block", alpha: float = 1.0, beta: float = 0.0, cols=None):
        check_op(lib().rcg_operator_apply(self.h, X.h, Y.h, alpha, beta,
                                          X.cols if cols is None else cols), self)

    def close(self):
        if getattr(self, "h", None):
            lib().rcg_operator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DenseOperator:
    """Dense real symmetric operator on the device (the user Otype as a dense
    matrix); its filter steps are MFMA block products."""

    def __init__(self, A, ctx: Context | None = None, device_ptr: int | None = None, n=None):
        self.ctx = ctx or default_context()
        self.is_complex = False
        self.dtype = np.dtype(np.float64)
        h = C.c_void_p()
        if device_ptr is not None:  # A already in HBM (column-major, lda = n)
            self.n = int(n)
            check(lib().rcg_operator_dense_create(self.ctx.h, self.n, C.c_void_p(device_ptr), self.n,
                                                  capi.MEM_DEVICE, C.byref(h)))
        else:
            A = np.asfortranarray(A, dtype=np.float64)
            if A.ndim != 2 or A.shape[0] != A.shape[1]:
                raise ValueError("a dense operator is a square matrix")
            self.n = A.shape[0]
            check(lib().rcg_operator_dense_create(self.ctx.h, self.n, A.ctypes.data_as(C.c_void_p),
                                                  self.n, capi.MEM_HOST, C.byref(h)))
        self.nnz = self.n * self.n
        self.h = h

    def apply(self, X: Block, Y: Block, alpha: float = 1.0, beta: float = 0.0, cols=None):
        check(lib().rcg_operator_apply(self.h, X.h, Y.h, alpha, beta, X.cols if cols is None else cols))

    def close(self):
        if getattr(self, "h", None):
            lib().rcg_operator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RandomizeOp:
    """Start-vector stream (ArandomizeOpClassLMCuda role, seeded)."""

    def __init__(self, seed: int = 0):
        self.seed = int(seed)
        self.counter = 0


class RayleighChebyshevLM:
    """RayleighChebyshevLM over librcg.so (RayleighChebyshevLM.h:213-472)."""

    def __init__(self):
        self.stopCondition = capi.STOP["COMBINATION"]
        self.maxMinTol = 1.0e-6
        self.maxInnerLoopCount = 10000
        self.fixedIterationCount = False
        self.nonRandomStartFlag = False
        self.verboseFlag = False
        self.minIntervalPolyDegreeMax = 200
        self.filterRepetitionCount = 1
        self.spectralBoundsIterMax = 10000
        self.filterPrecision = np.float64
        self.refine = False
        self._result = capi.SolveResult()
        self._residuals: list[float] = []
        self.trace = np.zeros((0, 10))

    # ---- setters (RayleighChebyshevLM.h:281-440)
    def setStopCondition(self, cond):
        if isinstance(cond, str):
            key = cond.upper()
            if key not in capi.STOP:
                raise RuntimeError(
                    "RayleighChebyshevLM Error : Stopping condition type specified not one of "
                    f"DEFAULT, COMBINATION, EIGENVALUE_ONLY,or RESIDUAL_ONLY. Offending specification : {cond}")
            cond = capi.STOP[key]
        self.stopCondition = capi.STOP["COMBINATION"] if cond == 0 else int(cond)

    def setMaxMinTolerance(self, val=1.0e-6):
        self.maxMinTol = float(val)

    def setMaxInnerLoopCount(self, val):
        self.maxInnerLoopCount = int(val)

    def resetMaxInnerLoopCount(self):
        self.maxInnerLoopCount = 10000

    def setFixedIterationCount(self, val):
        self.fixedIterationCount = bool(val)

    def setNonRandomStartFlag(self, val=True):
        self.nonRandomStartFlag = bool(val)

    def clearNonRandomStartFlag(self):
        self.nonRandomStartFlag = False

    def setVerboseFlag(self, val=True):
        self.verboseFlag = bool(val)

    def setMinIntervalPolyDegreeMax(self, val):
        self.minIntervalPolyDegreeMax = int(val)

    def setFilterRepetitionCount(self, val):
        self.filterRepetitionCount = int(val)

    def setMaxSpectralBoundsIter(self, val):
        self.spectralBoundsIterMax = int(val)

    # ---- getters
    def getEigVectorResiduals(self):
        return list(self._residuals)

    def getCountData(self):
        r = self._result
        return {"ortho": r.count_ortho, "OpApply": r.count_apply, "eigenvalue": r.count_eig}

    def getFinalData(self):
        r = self._result
        return {"maxResidual": r.final_maxres, "maxEigValueDifference": r.final_maxdiff,
                "maxRelEigValueGap": r.final_maxgap}

    def getSpectralBoundsIterCount(self):
        return self._result.lanczos_iters

    def getMinEigValueEst(self):
        return self._result.min_est

    def getMaxEigValueEst(self):
        return self._result.max_est

    def getGuardEigenvalue(self):
        return self._result.guard

    def getResult(self) -> capi.SolveResult:
        return self._result

    # ---- solves
    def _solve(self, k, tol, inc, buf, oP, randOp, eigValues, eigVectors, interval=False,
               lambda_max=0.0, bounds=None):
        p = capi.SolveParams(
            k=int(k), inc=int(inc), buf=int(buf), tol=float(tol), seed=randOp.seed,
            stop=self.stopCondition, interval=int(interval),
            max_inner=self.maxInnerLoopCount, deg_max=self.minIntervalPolyDegreeMax,
            rep=self.filterRepetitionCount, lambda_max=float(lambda_max),
            lanczos_itmax=self.spectralBoundsIterMax, maxmin_tol=self.maxMinTol,
            fixed_iteration=int(self.fixedIterationCount),
            filter_dtype=capi.F64 if self.filterPrecision == np.float64 else capi.F32,
            refine=int(self.refine), verbose=int(self.verboseFlag),
            use_bounds=0 if bounds is None else 1,
            nonrandom_start=(min(getattr(eigVectors, "valid_cols", 0), eigVectors.cols)
                             if (self.nonRandomStartFlag and eigVectors is not None) else 0),
            min_eig=0.0 if bounds is None else float(bounds[0]),
            max_eig=0.0 if bounds is None else float(bounds[1]))
        cap = max(int(k), int(inc) + int(buf), 1)
        ev = np.zeros(cap)
        rr = np.zeros(cap)
        trace = np.zeros((200000, 10))
        r = capi.SolveResult()
        check_op(lib().rcg_solve_lowest(
            oP.h, C.byref(p), C.byref(r), ev.ctypes.data_as(C.c_void_p),
            rr.ctypes.data_as(C.c_void_p), cap, eigVectors.h if eigVectors is not None else None,
            trace.ctypes.data_as(C.c_void_p), trace.shape[0]), oP)
        self._result = r
        f = int(r.found)
        self._residuals = rr[:f].tolist()
        self.trace = trace[: min(int(r.trace_len), trace.shape[0])].copy()
        randOp.counter = int(r.rng_counter)
        if eigVectors is not None:
            eigVectors.valid_cols = min(f, eigVectors.cols)
        if eigValues is not None:
            eigValues.clear()
            eigValues.extend(ev[:f].tolist())
        return f

    def getMinEigenSystem(self, eigCount, *args):
        """getMinEigenSystem(eigCount, subspaceTol, inc, buf, vStart, oP, randOp,
        eigValues, eigVectors) or getMinEigenSystem(eigCount, minEigValueEst,
        maxEigValueBound, subspaceTol, inc, buf, vStart, oP, randOp, eigValues,
        eigVectors) (RayleighChebyshevLM.h:693-738).  eigVectors is a Block
        with at least eigCount columns (or None)."""
        if len(args) == 8:
            tol, inc, buf, _vStart, oP, randOp, eigValues, eigVectors = args
            return self._solve(eigCount, tol, inc, buf, oP, randOp, eigValues, eigVectors)
        if len(args) == 10:
            mn, mx, tol, inc, buf, _vStart, oP, randOp, eigValues, eigVectors = args
            return self._solve(eigCount, tol, inc, buf, oP, randOp, eigValues, eigVectors,
                               bounds=(mn, mx))
        raise TypeError("getMinEigenSystem takes 9 or 11 arguments")

    def getMinIntervalEigenSystem(self, *args):
        """All eigenpairs below lambdaMax, at most maxEigensystemDim:
        getMinIntervalEigenSystem(lambdaMax, subspaceTol, inc, buf, maxEigensystemDim,
        vStart, oP, randOp, eigValues, eigVectors) (RayleighChebyshevLM.h:745-782) or,
        with the caller's spectral bounds and no Lanczos,
        getMinIntervalEigenSystem(minEigValueEst, lambdaMax, maxEigValueBound,
        subspaceTol, inc, buf, maxEigensystemDim, vStart, oP, randOp, eigValues,
        eigVectors) (RayleighChebyshev.h:812-821)."""
        if len(args) == 10:
            lambdaMax, tol, inc, buf, dim, _v, oP, randOp, eigValues, eigVectors = args
            return self._solve(dim, tol, inc, buf, oP, randOp, eigValues, eigVectors,
                               interval=True, lambda_max=lambdaMax)
        if len(args) == 12:
            mn, lambdaMax, mx, tol, inc, buf, dim, _v, oP, randOp, eigValues, eigVectors = args
            return self._solve(dim, tol, inc, buf, oP, randOp, eigValues, eigVectors,
                               interval=True, lambda_max=lambdaMax, bounds=(mn, mx))
        raise TypeError("getMinIntervalEigenSystem takes 10 or 12 arguments")


class RayleighChebyshev(RayleighChebyshevLM):
    """The older RayleighChebyshev template's calling convention
    (RayleighChebyshev.h:210-821): eigenvectors come back as a list of host
    vectors (numpy arrays) instead of a device block, and with the
    non-random-start flag the list's vectors are the start block.  The solve is
    the same device solve."""

    def _host_solve(self, call, k, inc, buf, oP, eigVectors):
        n = oP.n
        cols = max(int(k), int(inc) + int(buf))
        V = Block(n, cols, dtype=getattr(oP, "dtype", np.float64), ctx=oP.ctx)
        if self.nonRandomStartFlag and eigVectors:
            start = np.stack([np.asarray(v) for v in eigVectors[:cols]], axis=1)
            V.upload(start)
            V.valid_cols = start.shape[1]
        f = call(V)
        if eigVectors is not None:
            X = V.download(f)
            eigVectors.clear()
            eigVectors.extend(X[:, j].copy() for j in range(f))
        return f

    def getMinEigenSystem(self, eigCount, *args):
        if len(args) == 8:
            tol, inc, buf, _v, oP, randOp, eigValues, eigVectors = args
            return self._host_solve(lambda V: self._solve(eigCount, tol, inc, buf, oP, randOp,
                                                          eigValues, V),
                                    eigCount, inc, buf, oP, eigVectors)
        if len(args) == 10:
            mn, mx, tol, inc, buf, _v, oP, randOp, eigValues, eigVectors = args
            return self._host_solve(lambda V: self._solve(eigCount, tol, inc, buf, oP, randOp,
                                                          eigValues, V, bounds=(mn, mx)),
                                    eigCount, inc, buf, oP, eigVectors)
        raise TypeError("getMinEigenSystem takes 9 or 11 arguments")

    def getMinIntervalEigenSystem(self, *args):
        if len(args) == 10:
            lambdaMax, tol, inc, buf, dim, _v, oP, randOp, eigValues, eigVectors = args
            bounds = None
        elif len(args) == 12:
            mn, lambdaMax, mx, tol, inc, buf, dim, _v, oP, randOp, eigValues, eigVectors = args
            bounds = (mn, mx)
        else:
            raise TypeError("getMinIntervalEigenSystem takes 10 or 12 arguments")
        return self._host_solve(lambda V: self._solve(dim, tol, inc, buf, oP, randOp, eigValues, V,
                                                      interval=True, lambda_max=lambdaMax,
                                                      bounds=bounds),
                                dim, inc, buf, oP, eigVectors)
