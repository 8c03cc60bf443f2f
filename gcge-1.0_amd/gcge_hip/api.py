"""High-level Python access to the gfx950 GCGE backend, through its C ABI.

``solve`` drives the same calls a C program would make: HIP_MatCreateFromCSR,
GCGE_HIP_Solver_Init (options as -gcge_* arguments), HIP_MultiVecSetFromHost
for the initial block, GCGE_SOLVER_Solve or GCGE_HIP_SOLVER_SolveMixed, then
reads eval/evec back.  ``Ops`` calls individual GCGE_OPS slots (the device
kernels) on device multivectors, for kernel-level tests and measurements.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import capi
from .capi import DP, IP

# options parsed by GCGE_PARA_SetFromCommandLine (set before the workspace is sized)
_CLI = {"ev_max_it", "block_size", "given_init_evec", "ev_tol", "conv_type", "conv_omega_norm", "orth_type",
        "w_orth_type", "p_orth_type", "x_orth_type", "orth_zero_tol", "reorth_tol", "scbgs_reorth_tol",
        "max_reorth_time", "max_direct_orth_length", "multi_tol", "multi_tol_for_lock", "if_use_bcg",
        "cg_max_it", "cg_rate", "cg_type", "print_eval", "print_part_time", "print_para", "print_result",
        "opt_rr_eig_partly", "x_orth_block_size", "w_orth_block_size", "if_use_cg", "print_cg_error",
        "print_orth_zero", "scbgs_wself_max_reorth_time", "bcg_precision", "lock_converged", "bcg_precond",
        "rr_eig_full"}
# members set directly after initialisation (read at solve time)
_DIRECT = {"dirichlet_boundary", "stagnation_window", "stagnation_ratio", "stagnation_start"}


def _dp(a):
    return a.ctypes.data_as(DP)


def _ip(a):
    return a.ctypes.data_as(IP)


def _check(what):
    capi.raise_on_error(what)


def locality_order(M, leaf=0):
    """HIP_MatLocalityOrder of a CSR pattern: perm[new position] = row."""
    rp = np.ascontiguousarray(M.indptr, dtype=np.int32)
    ci = np.ascontiguousarray(M.indices, dtype=np.int32)
    perm = np.zeros(M.shape[0], dtype=np.int32)
    rc = capi.load().HIP_MatLocalityOrder(M.shape[0], _ip(rp), _ip(ci), leaf, _ip(perm))
    if rc != 0:
        raise ValueError(f"HIP_MatLocalityOrder: bad pattern ({rc})")
    return perm


class Matrix:
    """A CSR matrix uploaded to the device (HIP_MAT).  order: None (the
    caller's row order), "locality" (HIP_MatLocalityOrder) or a permutation
    perm[new position] = row (HIP_MatCreateFromCSROrdered); multivectors
    created from the matrix then transfer in the caller's order."""

    def __init__(self, M, order=None):
        L = capi.load()
        self.n = M.shape[0]
        self.rowptr = np.ascontiguousarray(M.indptr, dtype=np.int32)
        self.col = np.ascontiguousarray(M.indices, dtype=np.int32)
        self.val = np.ascontiguousarray(M.data, dtype=np.float64)
        self.nnz = int(self.col.size)
        if order is None:
            self.ptr = L.HIP_MatCreateFromCSR(self.n, M.shape[1], self.nnz, _ip(self.rowptr), _ip(self.col),
                                              _dp(self.val))
        else:
            perm = None if isinstance(order, str) and order == "locality" else \
                np.ascontiguousarray(order, dtype=np.int32)
            self.ptr = L.HIP_MatCreateFromCSROrdered(self.n, M.shape[1], self.nnz, _ip(self.rowptr), _ip(self.col),
                                                     _dp(self.val), _ip(perm) if perm is not None else None)
        if not self.ptr:
            _check("HIP_MatCreateFromCSR")
            raise RuntimeError("HIP_MatCreateFromCSR failed")

    def order(self):
        """The row order (perm[new position] = row), or None."""
        perm = np.zeros(self.n, dtype=np.int32)
        return perm if capi.load().HIP_MatGetOrder(self.ptr, _ip(perm)) == 1 else None

    @classmethod
    def from_file(cls, path):
        """HIP_ReadMatFile: the reference's CSR text format or the binary form."""
        L = capi.load()
        ptr = L.HIP_ReadMatFile(str(path).encode())
        if not ptr:
            _check("HIP_ReadMatFile")
            raise RuntimeError("HIP_ReadMatFile failed")
        self = cls.__new__(cls)
        self.ptr = ptr
        self.n = L.HIP_MatNumRows(ptr)
        self.nnz = L.HIP_MatNumNonzeros(ptr)
        return self

    def set_values(self, val):
        v = np.ascontiguousarray(val, dtype=np.float64)
        capi.load().HIP_MatSetValues(self.ptr, _dp(v))
        _check("HIP_MatSetValues")

    def set_precision(self, prec):
        capi.load().HIP_MatSetPrecision(self.ptr, prec)
        _check("HIP_MatSetPrecision")

    def close(self):
        if self.ptr:
            p = C.c_void_p(self.ptr)
            capi.load().HIP_MatDestroy(C.byref(p))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiVec:
    """A device multivector (n x ncols, column-major) created from a Matrix."""

    def __init__(self, mat: Matrix, ncols: int, prec: int = capi.GCGE_HIP_FP64, data=None):
        L = capi.load()
        mat.set_precision(prec)
        h = C.c_void_p()
        L.HIP_MultiVecCreate(C.byref(h), ncols, mat.ptr)
        mat.set_precision(capi.GCGE_HIP_FP64)
        _check("HIP_MultiVecCreate")
        self.ptr = h.value
        self.n = mat.n
        self.ncols = ncols
        self.prec = prec
        if data is not None:
            self.set(data)

    def set(self, X, start=0):
        X = np.asfortranarray(X, dtype=np.float64)
        if X.ndim == 1:
            X = X[:, None]
        capi.load().HIP_MultiVecSetFromHost(self.ptr, start, X.shape[1], _dp(X), X.shape[0])
        _check("HIP_MultiVecSetFromHost")

    def get(self, start=0, ncols=None):
        ncols = self.ncols - start if ncols is None else ncols
        X = np.zeros((self.n, ncols), order="F")
        capi.load().HIP_MultiVecGetToHost(self.ptr, start, ncols, _dp(X), self.n)
        _check("HIP_MultiVecGetToHost")
        return X

    def close(self):
        if self.ptr:
            p = C.c_void_p(self.ptr)
            capi.load().HIP_MultiVecDestroy(C.byref(p))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_matrix_file(path, M, binary=False):
    """HIP_WriteMatFile of a scipy CSR matrix (text: the reference's format)."""
    rp = np.ascontiguousarray(M.indptr, dtype=np.int32)
    ci = np.ascontiguousarray(M.indices, dtype=np.int32)
    v = np.ascontiguousarray(M.data, dtype=np.float64)
    capi.load().HIP_WriteMatFile(str(path).encode(), M.shape[0], M.shape[1], int(ci.size), _ip(rp), _ip(ci), _dp(v),
                                 1 if binary else 0)
    _check("HIP_WriteMatFile")


def _r(a, b):
    return (C.c_int * 2)(*a), (C.c_int * 2)(*b)


class Ops:
    """Calls GCGE_OPS slots of the HIP backend on device multivectors."""

    def __init__(self):
        L = capi.load()
        self._p = C.POINTER(capi.GCGE_OPS)()
        L.GCGE_OPS_Create(C.byref(self._p))
        L.GCGE_HIP_SetOps(self._p)
        self.o = self._p.contents

    def _f(self, name, ftype):
        return C.cast(getattr(self.o, name), ftype)

    def matdot(self, A: Matrix, X: MultiVec, Y: MultiVec, xs=0, ys=0, k=None):
        k = X.ncols - xs if k is None else k
        s, e = _r((xs, ys), (xs + k, ys + k))
        self._f("MatDotMultiVec", capi.F_MatDotMultiVec)(A.ptr, X.ptr, Y.ptr, s, e, self._p)
        _check("MatDotMultiVec")

    def matdot_dot(self, A: Matrix, X: MultiVec, Y: MultiVec, k=None):
        k = X.ncols if k is None else k
        out = np.zeros(k)
        s, e = _r((0, 0), (k, k))
        self._f("MatDotMultiVecInnerProdDiag", capi.F_MatDotMultiVecInnerProdDiag)(A.ptr, X.ptr, Y.ptr, s, e,
                                                                                    _dp(out), self._p)
        _check("MatDotMultiVecInnerProdDiag")
        return out

    def axpby(self, a, X: MultiVec, b, Y: MultiVec, xs=0, ys=0, k=None):
        k = X.ncols - xs if k is None else k
        s, e = _r((xs, ys), (xs + k, ys + k))
        self._f("MultiVecAxpby", capi.F_MultiVecAxpby)(a, X.ptr, b, Y.ptr, s, e, self._p)
        _check("MultiVecAxpby")

    def axpby_diag(self, a, X: MultiVec, b, Y: MultiVec):
        k = len(a)
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        s, e = _r((0, 0), (k, k))
        self._f("MultiVecAxpbyDiag", capi.F_MultiVecAxpbyDiag)(_dp(a), X.ptr, _dp(b), Y.ptr, s, e, self._p)
        _check("MultiVecAxpbyDiag")

    def dots(self, X: MultiVec, Y: MultiVec, k=None):
        k = X.ncols if k is None else k
        out = np.zeros(k)
        s, e = _r((0, 0), (k, k))
        self._f("MultiVecInnerProdDiag", capi.F_MultiVecInnerProdDiag)(X.ptr, Y.ptr, _dp(out), s, e, self._p)
        _check("MultiVecInnerProdDiag")
        return out

    def gram(self, V: MultiVec, W: MultiVec, vs=0, ve=None, ws=0, we=None, sym=False):
        ve = V.ncols if ve is None else ve
        we = W.ncols if we is None else we
        p = (we - ws) if sym else (ve - vs)
        q = we - ws
        a = np.zeros((p, q), order="F")
        s, e = _r((vs, ws), (ve, we))
        self._f("MultiVecInnerProd", capi.F_MultiVecInnerProd)(V.ptr, W.ptr, _dp(a), b"sym" if sym else b"ns", s, e,
                                                               p, self._p)
        _check("MultiVecInnerProd")
        return a

    def lincomb(self, X: MultiVec, Y: MultiVec, coef, xs=0, ys=0, alpha=1.0, beta=0.0):
        coef = np.asfortranarray(coef, dtype=np.float64)
        p, q = coef.shape
        s, e = _r((xs, ys), (xs + p, ys + q))
        if alpha == 1.0 and beta == 0.0:
            self._f("MultiVecLinearComb", capi.F_MultiVecLinearComb)(X.ptr, Y.ptr, s, e, _dp(coef), p, None, 0,
                                                                     self._p)
        else:
            self._f("MultiVecLinearCombAxpby", capi.F_MultiVecLinearCombAxpby)(X.ptr, Y.ptr, s, e, _dp(coef), p,
                                                                               alpha, beta, self._p)
        _check("MultiVecLinearComb")

    def swap(self, X: MultiVec, Y: MultiVec, xs, ys, k):
        s, e = _r((xs, ys), (xs + k, ys + k))
        self._f("MultiVecSwap", capi.F_MultiVecSwap)(X.ptr, Y.ptr, s, e, self._p)
        _check("MultiVecSwap")

    def copy(self, X: MultiVec, Y: MultiVec, xs=0, ys=0, k=None):
        k = X.ncols - xs if k is None else k
        s, e = _r((xs, ys), (xs + k, ys + k))
        self._f("MultiVecCopy", capi.F_MultiVecCopy)(X.ptr, Y.ptr, s, e, self._p)
        _check("MultiVecCopy")

    def gather(self, X: MultiVec, idx, Y: MultiVec, ys=0):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        self._f("MultiVecGather", capi.F_MultiVecGather)(X.ptr, _ip(idx), len(idx), Y.ptr, ys, self._p)
        _check("MultiVecGather")

    def cg_update(self, X: MultiVec, xs, P: MultiVec, W: MultiVec, R: MultiVec, k, alpha):
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        rr = np.zeros(k)
        self._f("BlockCGUpdate", capi.F_BlockCGUpdate)(X.ptr, xs, P.ptr, W.ptr, R.ptr, k, _dp(alpha), _dp(rr),
                                                       self._p)
        _check("BlockCGUpdate")
        return rr

    def cg_residual_update(self, W: MultiVec, R: MultiVec, k, alpha):
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        rr = np.zeros(k)
        self._f("BlockCGResidualUpdate", capi.F_BlockCGResidualUpdate)(W.ptr, R.ptr, k, _dp(alpha), _dp(rr),
                                                                       self._p)
        _check("BlockCGResidualUpdate")
        return rr

    def pcg_residual_update(self, A: Matrix, W: MultiVec, R: MultiVec, k, alpha):
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        rz, rr = np.zeros(k), np.zeros(k)
        self._f("BlockPCGResidualUpdate", capi.F_BlockPCGResidualUpdate)(A.ptr, W.ptr, R.ptr, k, _dp(alpha),
                                                                         _dp(rz), _dp(rr), self._p)
        _check("BlockPCGResidualUpdate")
        return rz, rr

    def pcg_search_update(self, A: Matrix, X: MultiVec, xs, P: MultiVec, R: MultiVec, kx, kp, alpha, beta):
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        self._f("BlockPCGSearchUpdate", capi.F_BlockPCGSearchUpdate)(A.ptr, X.ptr, xs, P.ptr, R.ptr, kx, kp,
                                                                     _dp(alpha), _dp(beta), self._p)
        _check("BlockPCGSearchUpdate")

    def cg_search_update(self, X: MultiVec, xs, P: MultiVec, R: MultiVec, kx, kp, alpha, beta):
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        self._f("BlockCGSearchUpdate", capi.F_BlockCGSearchUpdate)(X.ptr, xs, P.ptr, R.ptr, kx, kp, _dp(alpha),
                                                                   _dp(beta), self._p)
        _check("BlockCGSearchUpdate")

    def residual_norms(self, A: Matrix, B, X: MultiVec, idx, lam):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        k = len(idx)
        rn, xn, bn = np.zeros(k), np.zeros(k), np.zeros(k)
        self._f("MultiVecResidualNorms", capi.F_MultiVecResidualNorms)(
            A.ptr, B.ptr if B is not None else None, X.ptr, _ip(idx), k, _dp(lam), _dp(rn), _dp(xn), _dp(bn),
            self._p)
        _check("MultiVecResidualNorms")
        return rn, xn, bn

    def dirichlet(self, X: MultiVec, nev, A: Matrix, B: Matrix):
        self._f("SetDirichletBoundary", capi.F_SetDirichletBoundary)(X.ptr, nev, A.ptr, B.ptr)
        _check("SetDirichletBoundary")

    def close(self):
        if self._p:
            capi.load().GCGE_OPS_Free(C.byref(self._p))
            self._p = None


def synchronize():
    capi.load().GCGE_HIP_Synchronize()
    _check("synchronize")


# GCGE_HIP_MIXED_OPTS of the two-phase forms (0 selects the library's default for
# fp32_tol, the iteration budgets -- the caller's ev_max_it -- and stagnation_ratio)
MIXED_DEFAULTS = dict(fp32_tol=8e-3, fp32_max_it=0, fp64_max_it=0, stagnation_window=2, stagnation_ratio=0.5,
                      fp32_w_orth=1, refine_w_orth=1, refine_bcg_precision=1, stagnation_start=1e-3, cg_max_it=30,
                      bcg_precond=1, rr_eig_full=1)
FP32_REFINE_DEFAULTS = dict(MIXED_DEFAULTS, fp32_w_orth=0, refine_w_orth=0, refine_bcg_precision=0, cg_max_it=0,
                            bcg_precond=0, rr_eig_full=0)


class Solver:
    """A set-up GCGE solver on the device, reusable across solves.

    GCGE_HIP_Solver_Init runs once (workspace allocation included); each
    ``solve(X0)`` writes the initial block and calls GCGE_SOLVER_Solve or
    GCGE_HIP_SOLVER_SolveMixed, which restart the per-solve state and reuse
    the workspace (and, for the two-phase forms, the FP32 phase's solver).
    ``given_x0`` says whether the solves receive an initial block
    (-gcge_given_init_evec 1) or draw it with rand()."""

    def __init__(self, A, B, nev, precision="fp64", mixed=None, seed=None, given_x0=True, **params):
        L = capi.load()
        self._own = []
        self.mA = A if isinstance(A, Matrix) else Matrix(A)
        if not isinstance(A, Matrix):
            self._own.append(self.mA)
        self.mB = None if B is None else (B if isinstance(B, Matrix) else Matrix(B))
        if B is not None and not isinstance(B, Matrix):
            self._own.append(self.mB)
        if precision not in ("fp64", "fp64-cg32", "mixed", "fp32+refine"):
            raise ValueError(precision)
        self.precision = precision
        self.nev = nev
        self.n = self.mA.n
        self.params = params
        self.seed = seed
        args = [b"gcge"]
        for k, v in params.items():
            if k in _CLI:
                args += [("-gcge_" + k).encode(), str(v).encode()]
        if given_x0:
            args += [b"-gcge_given_init_evec", b"1"]
        args += [b"-gcge_print_para", str(params.get("print_para", 0)).encode(),
                 b"-gcge_print_result", str(params.get("print_result", 0)).encode()]
        self._argv = (C.c_char_p * len(args))(*args)
        if seed is not None:
            self._srand(seed)
        self.solver = L.GCGE_HIP_Solver_Init(self.mA.ptr, self.mB.ptr if self.mB else None, nev, len(args),
                                             self._argv)
        if not self.solver:
            _check("GCGE_HIP_Solver_Init")
            raise RuntimeError("GCGE_HIP_Solver_Init failed")
        sp = self.solver.contents
        if sp.workspace.contents.V is None:
            self.close()
            raise RuntimeError("solver setup failed (missing operations)")
        para = sp.para.contents
        for k in _DIRECT:
            if k in params:
                setattr(para, k, params[k])
        if precision == "fp64-cg32":
            para.bcg_precision = 1
        self.mixed_opts = None
        if precision in ("mixed", "fp32+refine"):
            o = capi.GCGE_HIP_MIXED_OPTS()
            opts = dict(MIXED_DEFAULTS if precision == "mixed" else FP32_REFINE_DEFAULTS)
            opts.update(mixed or {})
            for f, _ in capi.GCGE_HIP_MIXED_OPTS._fields_:
                setattr(o, f, opts[f])
            self.mixed_opts = o

    @staticmethod
    def _srand(seed):
        import ctypes.util
        libc = C.CDLL(ctypes.util.find_library("c"))
        libc.srand(C.c_uint(seed))

    def set_x0(self, X0):
        """Write the initial block: an (n, nev) host array or an FP64 MultiVec in HBM."""
        L = capi.load()
        sp = self.solver.contents
        if isinstance(X0, MultiVec):
            # a block already in HBM: device-to-device copy into evec
            s, e = _r((0, 0), (self.nev, self.nev))
            fn = C.cast(sp.ops.contents.MultiVecCopy, capi.F_MultiVecCopy)
            fn(X0.ptr, sp.evec, s, e, sp.ops)
        else:
            X0 = np.asfortranarray(X0, dtype=np.float64)
            L.HIP_MultiVecSetFromHost(sp.evec, 0, self.nev, _dp(X0), self.n)
        _check("initial block")

    def solve(self, X0=None, keep_evec=True):
        L = capi.load()
        sp = self.solver.contents
        para = sp.para.contents
        nev, n = self.nev, self.n
        if X0 is not None:
            self.set_x0(X0)
        info = {}
        rc = 0
        t0 = time.perf_counter()
        if self.mixed_opts is None:
            L.GCGE_SOLVER_Solve(self.solver)
        else:
            mi = capi.GCGE_HIP_MIXED_INFO()
            rc = L.GCGE_HIP_SOLVER_SolveMixed(self.solver, C.byref(self.mixed_opts), C.byref(mi))
            info = {f: getattr(mi, f) for f, _ in capi.GCGE_HIP_MIXED_INFO._fields_}
        L.GCGE_HIP_Synchronize()
        wall = time.perf_counter() - t0
        _check("solve")
        ev = np.array(sp.eval[:nev])
        res = np.array(para.res[:nev])
        X = None
        if keep_evec:
            X = np.zeros((n, nev), order="F")
            L.HIP_MultiVecGetToHost(sp.evec, 0, nev, _dp(X), n)
            _check("download evec")
        out = dict(eval=ev, evec=X, res=res, num_iter=para.num_iter, num_unlock=para.num_unlock,
                   seconds=wall if self.mixed_opts is not None else para.solve_seconds, wall=wall,
                   bcg_calls=para.bcg_calls, bcg_iters_total=para.bcg_iters_total, info=info, rc=rc)
        pt = para.stat_para.contents.part_time_total.contents
        out["part_time"] = {f: getattr(pt, f) for f, _ in capi.GCGE_PART_TIME._fields_}
        return out

    def residual_norms(self):
        """||A x_i - lambda_i B x_i||, ||x_i|| and ||B x_i|| of the solver's
        current eigenpairs, in FP64 on the device (MultiVecResidualNorms)."""
        sp = self.solver.contents
        nev = self.nev
        idx = np.arange(nev, dtype=np.int32)
        lam = np.array(sp.eval[:nev])
        rn, xn, bn = np.zeros(nev), np.zeros(nev), np.zeros(nev)
        fn = C.cast(sp.ops.contents.MultiVecResidualNorms, capi.F_MultiVecResidualNorms)
        fn(self.mA.ptr, self.mB.ptr if self.mB else None, sp.evec, _ip(idx), nev, _dp(lam), _dp(rn), _dp(xn),
           _dp(bn), sp.ops)
        _check("MultiVecResidualNorms")
        return rn, xn, bn

    def close(self):
        if getattr(self, "solver", None):
            capi.load().GCGE_SOLVER_Free_All(C.byref(self.solver))
            self.solver = None
        for m in getattr(self, "_own", []):
            m.close()
        self._own = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve(A, B, nev, X0=None, precision="fp64", mixed=None, seed=None, keep_evec=True, **params):
    """Solve A x = lambda B x on the device.

    precision:
      "fp64"         everything FP64 (the drop-in);
      "fp64-cg32"    FP64, with the block-CG inner solves in FP32 as a
                     correction of the FP64 residual (para.bcg_precision = 1);
      "mixed"        GCGE_HIP_SOLVER_SolveMixed: an FP32 phase (FP32 matrix
                     values and multivectors, CholQR W orthonormalization),
                     then FP64 refinement sweeps from its eigenvectors with
                     FP32 inner solves; inner solves preconditioned with the
                     diagonal of A and capped at 30 CG steps in both phases
                     (mixed=dict(cg_max_it=0, bcg_precond=0) keeps the
                     caller's settings);
      "fp32+refine"  the same two phases with the reference's own
                     orthonormalization and FP64 inner solves in the
                     refinement.
    X0 is an (n, nev) host array or an FP64 MultiVec already on the device.
    ``mixed`` overrides GCGE_HIP_MIXED_OPTS fields.  Returns a dict with
    eval, evec, res, num_iter, num_unlock, seconds (GCGE_Solve wall time, or
    the whole two-phase solve), bcg statistics, the return code ``rc`` of the
    two-phase call and, for the two-phase forms, the phase info."""
    s = Solver(A, B, nev, precision=precision, mixed=mixed, seed=seed, given_x0=X0 is not None, **params)
    try:
        return s.solve(X0, keep_evec=keep_evec)
    finally:
        s.close()


def eig_gradient_generalized(A: Matrix, B, X, lam, weights, nparam, jA, jB=None):
    """d(sum_i w_i lambda_i)/d theta for A(theta) x = lambda B(theta) x via
    GCGE_HIP_EigGradientGeneralized; X holds B-normalized eigenvectors, jA
    and jB are (jrowptr, jnz, jval) Jacobians of A's and B's stored values
    (jB None: B does not depend on theta)."""
    L = capi.load()
    k = X.shape[1]
    mv = MultiVec(A, k, data=X)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)

    def arrs(j):
        return (np.ascontiguousarray(j[0], dtype=np.int32), np.ascontiguousarray(j[1], dtype=np.int32),
                np.ascontiguousarray(j[2], dtype=np.float64))

    a = arrs(jA)
    b = arrs(jB) if jB is not None else None
    grad = np.zeros(nparam)
    L.GCGE_HIP_EigGradientGeneralized(A.ptr, B.ptr if B is not None else None, mv.ptr, k, _dp(lam), _dp(w), nparam,
                                      _ip(a[0]), _ip(a[1]), _dp(a[2]), _ip(b[0]) if b else None,
                                      _ip(b[1]) if b else None, _dp(b[2]) if b else None, _dp(grad))
    _check("GCGE_HIP_EigGradientGeneralized")
    mv.close()
    return grad


def eig_gradient(A: Matrix, X, weights, nparam, jrowptr, jnz, jval):
    """d(sum_i w_i lambda_i)/d theta via GCGE_HIP_EigGradient (B = I)."""
    L = capi.load()
    k = X.shape[1]
    mv = MultiVec(A, k, data=X)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    jrowptr = np.ascontiguousarray(jrowptr, dtype=np.int32)
    jnz = np.ascontiguousarray(jnz, dtype=np.int32)
    jval = np.ascontiguousarray(jval, dtype=np.float64)
    grad = np.zeros(nparam)
    L.GCGE_HIP_EigGradient(A.ptr, mv.ptr, k, _dp(w), nparam, _ip(jrowptr), _ip(jnz), _dp(jval), _dp(grad))
    _check("GCGE_HIP_EigGradient")
    mv.close()
    return grad
