"""ctypes binding of libgcge_hip.so (the C ABI declared in include/*.h).

The structs mirror include/gcge_para.h, include/gcge_solver.h and
include/gcge_ops.h member for member; ``abi_check()`` compares their sizes and
offsets with the library's own.  Loading fails loudly when the library has not
been built: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libgcge_hip.so")

c_int = C.c_int
c_double = C.c_double
c_void_p = C.c_void_p
c_char_p = C.c_char_p
IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)

GCGE_HIP_FP64 = 0
GCGE_HIP_FP32 = 1


class GCGE_PART_TIME(C.Structure):
    _fields_ = [(n, c_double) for n in ("x_axpy_time", "x_orth_time", "p_axpy_time", "p_orth_time",
                                        "w_line_time", "w_orth_time", "rr_mat_time", "rr_eigen_time",
                                        "conv_time")]


class GCGE_STATISTIC_PARA(C.Structure):
    _fields_ = [("part_time_total", C.POINTER(GCGE_PART_TIME)),
                ("part_time_one_iter", C.POINTER(GCGE_PART_TIME)),
                ("ops_time_count_total", c_void_p), ("ops_time_count_linear_solver", c_void_p)]


class GCGE_ORTH_PARA(C.Structure):
    _fields_ = [("orth_zero_tol", c_double), ("reorth_tol", c_double), ("scbgs_reorth_tol", c_double),
                ("criterion_tol", c_double), ("max_reorth_time", c_int),
                ("scbgs_wself_max_reorth_time", c_int), ("print_orth_zero", c_int),
                ("x_orth_block_size", c_int), ("w_orth_block_size", c_int),
                ("max_direct_orth_length", c_int)]


class GCGE_PARA(C.Structure):
    _fields_ = [
        ("nev", c_int), ("ev_max_it", c_int), ("block_size", c_int),
        ("if_lobgcg", c_int), ("given_init_evec", c_int),
        ("ev_tol", c_double), ("conv_type", c_char_p), ("conv_omega_norm", c_double), ("num_unlock", c_int),
        ("orth_type", c_char_p), ("w_orth_type", c_char_p), ("p_orth_type", c_char_p), ("x_orth_type", c_char_p),
        ("orth_para", C.POINTER(GCGE_ORTH_PARA)),
        ("multi_tol", c_double), ("multi_tol_for_lock", c_double),
        ("if_use_cg", c_int), ("cg_max_it", c_int), ("cg_rate", c_double), ("cg_type", c_int),
        ("if_use_bcg", c_int),
        ("num_iter", c_int), ("max_res", c_double), ("max_ind", c_double), ("min_res", c_double),
        ("min_ind", c_double), ("sum_res", c_double), ("res", DP), ("dirichlet_boundary", c_int),
        ("use_mpi_bcast", c_int),
        ("print_eval", c_int), ("print_cg_error", c_int), ("print_orthzero", c_int), ("print_para", c_int),
        ("print_result", c_int), ("print_matlab", c_int), ("print_part_time", c_int), ("print_level", c_int),
        ("stat_para", C.POINTER(GCGE_STATISTIC_PARA)),
        ("opt_rr_eig_partly", c_int), ("opt_bcast", c_int), ("opt_allgatherv", c_int),
        ("bcg_iters_total", c_int), ("bcg_calls", c_int), ("spmv_columns", c_int),
        ("solve_seconds", c_double),
        ("stagnation_window", c_int), ("stagnation_ratio", c_double), ("stopped_on_stagnation", c_int),
        ("bcg_precision", c_int), ("stagnation_start", c_double), ("lock_converged", c_int),
        ("basis_eps", c_double), ("bcg_precond", c_int), ("rr_eig_full", c_int),
    ]


class GCGE_WORKSPACE(C.Structure):
    _fields_ = [
        ("V_size", c_int), ("V_tmp_size", c_int), ("CG_p_size", c_int), ("CG_r_size", c_int),
        ("RitzVec_size", c_int),
        ("V", c_void_p), ("V_tmp", c_void_p), ("RitzVec", c_void_p), ("evec", c_void_p), ("CG_p", c_void_p),
        ("CG_r", c_void_p),
        ("unlock", IP), ("max_dim_x", c_int), ("dim_x", c_int), ("dim_xp", c_int), ("dim_xpw", c_int),
        ("last_dim_xpw", c_int), ("last_dim_x", c_int), ("conv_bs", c_int), ("unconv_bs", c_int),
        ("num_soft_locked_in_last_iter", c_int), ("dense_matrix", c_void_p),
        ("subspace_matrix", DP), ("subspace_evec", DP), ("subspace_dtmp", DP), ("subspace_itmp", IP),
        ("eval", DP), ("orth_ind", IP), ("bcg32", c_void_p * 4), ("bcg32_cols", c_int),
        ("mixed32", c_void_p),
    ]


OPS_SLOTS = [
    "VecSetRandomValue", "MatDotVec", "VecAxpby", "VecInnerProd", "VecLocalInnerProd", "VecCreateByVec",
    "VecCreateByMat", "VecDestroy", "LinearSolver", "linear_solver_workspace", "Orthonormalization",
    "DenseMatCreate", "DenseMatDestroy", "MultiVecCreateByVec", "MultiVecCreateByMat",
    "MultiVecCreateByMultiVec", "MultiVecDestroy", "MultiVecSetRandomValue", "MatDotMultiVec", "MultiVecAxpby",
    "MultiVecAxpbyColumn", "MultiVecLinearComb", "MultiVecInnerProd", "MultiVecSwap", "MultiVecPrint",
    "GetVecFromMultiVec", "RestoreVecForMultiVec", "SetDirichletBoundary", "DenseMatEigenSolver",
    "DenseMatDotDenseMat", "DenseSymMatDotDenseMat", "ArrayDotArray", "ArrayNorm", "ArrayAXPBY", "ArrayCopy",
    "ArrayScale",
    # block-kernel slots of this build
    "MultiVecAxpbyDiag", "MultiVecInnerProdDiag", "MultiVecLinearCombAxpby", "MatDotMultiVecInnerProdDiag",
    "BlockCGUpdate", "MultiVecResidualNorms", "MultiVecGather", "MultiVecCopy", "MultiVecNumRows",
    "MultiVecCreateByMatPrecision", "BlockCGResidualUpdate", "BlockCGSearchUpdate", "BlockPCGResidualUpdate",
    "BlockPCGSearchUpdate",
]


class GCGE_OPS(C.Structure):
    _fields_ = [(n, c_void_p) for n in OPS_SLOTS]


class GCGE_SOLVER(C.Structure):
    _fields_ = [("A", c_void_p), ("B", c_void_p), ("evec", c_void_p), ("eval", DP),
                ("para", C.POINTER(GCGE_PARA)), ("ops", C.POINTER(GCGE_OPS)),
                ("workspace", C.POINTER(GCGE_WORKSPACE))]


class GCGE_HIP_MIXED_OPTS(C.Structure):
    _fields_ = [("fp32_tol", c_double), ("fp32_max_it", c_int), ("fp64_max_it", c_int),
                ("stagnation_window", c_int), ("stagnation_ratio", c_double), ("fp32_w_orth", c_int),
                ("refine_w_orth", c_int), ("refine_bcg_precision", c_int), ("stagnation_start", c_double),
                ("cg_max_it", c_int), ("bcg_precond", c_int), ("rr_eig_full", c_int)]


class GCGE_HIP_MIXED_INFO(C.Structure):
    _fields_ = [("fp32_iters", c_int), ("fp64_iters", c_int), ("fp32_seconds", c_double),
                ("fp64_seconds", c_double), ("fp64_num_unlock", c_int), ("refine_passes", c_int),
                ("fp32_failed", c_int)]

GCGE_HIP_NOT_CONVERGED = -2


SP = C.POINTER(GCGE_SOLVER)

# name -> (restype, argtypes)
SIGNATURES = {
    "GCGE_SOLVER_Create": (None, [C.POINTER(SP)]),
    "GCGE_SOLVER_Free": (None, [C.POINTER(SP)]),
    "GCGE_SOLVER_Free_All": (None, [C.POINTER(SP)]),
    "GCGE_SOLVER_SetMatA": (None, [SP, c_void_p]),
    "GCGE_SOLVER_SetMatB": (None, [SP, c_void_p]),
    "GCGE_SOLVER_SetEigenvalues": (None, [SP, DP]),
    "GCGE_SOLVER_SetEigenvectors": (None, [SP, c_void_p]),
    "GCGE_SOLVER_SetNumEigen": (None, [SP, c_int]),
    "GCGE_SOLVER_Setup": (None, [SP]),
    "GCGE_SOLVER_Solve": (None, [SP]),
    "GCGE_SOLVER_GetEigenvalues": (None, [SP, C.POINTER(DP)]),
    "GCGE_SOLVER_GetEigenvectors": (None, [SP, C.POINTER(c_void_p)]),
    "GCGE_SOLVER_SetOpsLinearSolverWorkspace": (None, [SP, c_void_p]),
    "GCGE_PARA_Create": (None, [C.POINTER(C.POINTER(GCGE_PARA))]),
    "GCGE_PARA_Free": (None, [C.POINTER(C.POINTER(GCGE_PARA))]),
    "GCGE_PARA_SetFromCommandLine": (c_int, [C.POINTER(GCGE_PARA), c_int, C.POINTER(c_char_p)]),
    "GCGE_PARA_SetNumEigen": (None, [C.POINTER(GCGE_PARA), c_int]),
    "GCGE_PARA_Setup": (None, [C.POINTER(GCGE_PARA)]),
    "GCGE_PrintParaInfo": (None, [C.POINTER(GCGE_PARA)]),
    "GCGE_PrintIterationInfo": (None, [DP, C.POINTER(GCGE_PARA)]),
    "GCGE_PrintFinalInfo": (None, [DP, C.POINTER(GCGE_PARA)]),
    "GCGE_STATISTIC_PARA_Create": (None, [c_void_p]),
    "GCGE_STATISTIC_PARA_Free": (None, [c_void_p]),
    "GCGE_OPS_Create": (None, [C.POINTER(C.POINTER(GCGE_OPS))]),
    "GCGE_OPS_Setup": (c_int, [C.POINTER(GCGE_OPS)]),
    "GCGE_OPS_Free": (None, [C.POINTER(C.POINTER(GCGE_OPS))]),
    "GCGE_OPS_SetLinearSolverWorkspace": (None, [C.POINTER(GCGE_OPS), c_void_p]),
    "GCGE_WORKSPACE_Create": (None, [c_void_p]),
    "GCGE_WORKSPACE_Setup": (None, [c_void_p, C.POINTER(GCGE_PARA), C.POINTER(GCGE_OPS), c_void_p]),
    "GCGE_WORKSPACE_Free": (None, [c_void_p, C.POINTER(GCGE_PARA), C.POINTER(GCGE_OPS)]),
    "GCGE_Solve": (None, [c_void_p, c_void_p, DP, c_void_p, C.POINTER(GCGE_PARA), C.POINTER(GCGE_OPS),
                          C.POINTER(GCGE_WORKSPACE)]),
    "GCGE_CheckConvergence": (None, [c_void_p, c_void_p, DP, c_void_p, C.POINTER(GCGE_OPS),
                                     C.POINTER(GCGE_PARA), C.POINTER(GCGE_WORKSPACE)]),
    "GCGE_CheckEvalMultiplicity": (None, [c_int, c_int, c_int, IP, DP]),
    "HIP_MatCreateFromCSR": (c_void_p, [c_int, c_int, c_int, IP, IP, DP]),
    "HIP_ReadMatFile": (c_void_p, [c_char_p]),
    "HIP_WriteMatFile": (c_int, [c_char_p, c_int, c_int, c_int, IP, IP, DP, c_int]),
    "HIP_MatLocalityOrder": (c_int, [c_int, IP, IP, c_int, IP]),
    "HIP_MatCreateFromCSROrdered": (c_void_p, [c_int, c_int, c_int, IP, IP, DP, IP]),
    "HIP_MatGetOrder": (c_int, [c_void_p, IP]),
    "HIP_MatGatherKind": (c_int, [c_void_p]),
    "HIP_MatDestroy": (None, [C.POINTER(c_void_p)]),
    "HIP_MatNumRows": (c_int, [c_void_p]),
    "HIP_MatNumNonzeros": (c_int, [c_void_p]),
    "HIP_MatSetValues": (c_int, [c_void_p, DP]),
    "HIP_MatSetPrecision": (c_int, [c_void_p, c_int]),
    "HIP_MultiVecCreate": (c_int, [C.POINTER(c_void_p), c_int, c_void_p]),
    "HIP_MultiVecDestroy": (None, [C.POINTER(c_void_p)]),
    "HIP_MultiVecSetFromHost": (c_int, [c_void_p, c_int, c_int, DP, c_int]),
    "HIP_MultiVecGetToHost": (c_int, [c_void_p, c_int, c_int, DP, c_int]),
    "HIP_MultiVecNumCols": (c_int, [c_void_p]),
    "HIP_MultiVecPrecision": (c_int, [c_void_p]),
    "GCGE_HIP_SetOps": (None, [C.POINTER(GCGE_OPS)]),
    "GCGE_SOLVER_SetHIPOps": (None, [SP]),
    "GCGE_HIP_Solver_Init": (SP, [c_void_p, c_void_p, c_int, c_int, C.POINTER(c_char_p)]),
    "GCGE_HIP_SOLVER_SolveMixed": (c_int, [SP, C.POINTER(GCGE_HIP_MIXED_OPTS), C.POINTER(GCGE_HIP_MIXED_INFO)]),
    "GCGE_HIP_EigGradient": (c_int, [c_void_p, c_void_p, c_int, DP, c_int, IP, IP, DP, DP]),
    "GCGE_HIP_EigGradientGeneralized": (c_int, [c_void_p, c_void_p, c_void_p, c_int, DP, DP, c_int, IP, IP, DP, IP,
                                                IP, DP, DP]),
    "GCGE_HIP_MultiVecInnerProdDiagNoAlloc": (c_int, [c_void_p, c_void_p, DP, IP, IP]),
    "GCGE_HIP_GetLastError": (c_int, [c_char_p, c_int]),
    "GCGE_HIP_ClearError": (None, []),
    "GCGE_HIP_Synchronize": (c_int, []),
    "GCGE_HIP_DeviceBytes": (c_double, []),
    "GCGE_HIP_AbiLayout": (c_int, [IP, c_int]),
    "GCGE_HIP_SetProfiling": (None, [c_int]),
    "GCGE_HIP_ResetKernelStats": (None, []),
    "GCGE_HIP_GetKernelStats": (c_int, [c_int, c_char_p, c_int, DP, DP, DP, DP]),
}

# CFUNCTYPE signatures of the ops slots called from Python
F_MatDotMultiVec = C.CFUNCTYPE(None, c_void_p, c_void_p, c_void_p, IP, IP, c_void_p)
F_MultiVecAxpby = C.CFUNCTYPE(None, c_double, c_void_p, c_double, c_void_p, IP, IP, c_void_p)
F_MultiVecLinearComb = C.CFUNCTYPE(None, c_void_p, c_void_p, IP, IP, DP, c_int, c_void_p, c_int, c_void_p)
F_MultiVecInnerProd = C.CFUNCTYPE(None, c_void_p, c_void_p, DP, c_char_p, IP, IP, c_int, c_void_p)
F_MultiVecSwap = C.CFUNCTYPE(None, c_void_p, c_void_p, IP, IP, c_void_p)
F_MultiVecAxpbyDiag = C.CFUNCTYPE(None, DP, c_void_p, DP, c_void_p, IP, IP, c_void_p)
F_MultiVecInnerProdDiag = C.CFUNCTYPE(None, c_void_p, c_void_p, DP, IP, IP, c_void_p)
F_MultiVecLinearCombAxpby = C.CFUNCTYPE(None, c_void_p, c_void_p, IP, IP, DP, c_int, c_double, c_double,
                                        c_void_p)
F_MatDotMultiVecInnerProdDiag = C.CFUNCTYPE(None, c_void_p, c_void_p, c_void_p, IP, IP, DP, c_void_p)
F_BlockCGUpdate = C.CFUNCTYPE(None, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, DP, DP, c_void_p)
F_BlockCGResidualUpdate = C.CFUNCTYPE(None, c_void_p, c_void_p, c_int, DP, DP, c_void_p)
F_BlockCGSearchUpdate = C.CFUNCTYPE(None, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, DP, DP, c_void_p)
F_BlockPCGResidualUpdate = C.CFUNCTYPE(None, c_void_p, c_void_p, c_void_p, c_int, DP, DP, DP, c_void_p)
F_BlockPCGSearchUpdate = C.CFUNCTYPE(None, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, DP, DP,
                                     c_void_p)
F_MultiVecResidualNorms = C.CFUNCTYPE(None, c_void_p, c_void_p, c_void_p, IP, c_int, DP, DP, DP, DP, c_void_p)
F_MultiVecGather = C.CFUNCTYPE(None, c_void_p, IP, c_int, c_void_p, c_int, c_void_p)
F_MultiVecCopy = C.CFUNCTYPE(None, c_void_p, c_void_p, IP, IP, c_void_p)
F_SetDirichletBoundary = C.CFUNCTYPE(None, c_void_p, c_int, c_void_p, c_void_p)

_lib = None


def load(path: str = LIB_PATH):
    """Load libgcge_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python gcge-1.0_amd/build.py` "
                           "(there is no fallback implementation)")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return list(SIGNATURES.keys())


def last_error():
    buf = C.create_string_buffer(512)
    code = load().GCGE_HIP_GetLastError(buf, 512)
    return code, buf.value.decode(errors="replace")


def raise_on_error(what=""):
    code, msg = last_error()
    if code:
        load().GCGE_HIP_ClearError()
        raise RuntimeError(f"gcge-hip error {code} {what}: {msg}")


def kernel_stats():
    """{class: dict(count, ms, bytes, flops)} from the backend profiler."""
    L = load()
    out = {}
    buf = C.create_string_buffer(64)
    cnt, ms, by, fl = c_double(), c_double(), c_double(), c_double()
    n = L.GCGE_HIP_GetKernelStats(0, buf, 64, C.byref(cnt), C.byref(ms), C.byref(by), C.byref(fl))
    for k in range(n):
        L.GCGE_HIP_GetKernelStats(k, buf, 64, C.byref(cnt), C.byref(ms), C.byref(by), C.byref(fl))
        out[buf.value.decode()] = dict(count=int(cnt.value), ms=ms.value, bytes=by.value, flops=fl.value)
    return out


def abi_check():
    """Compare struct sizes/offsets of this module with the library's."""
    L = load()
    out = (c_int * 32)()
    n = L.GCGE_HIP_AbiLayout(out, 32)
    mine = [C.sizeof(GCGE_PARA), GCGE_PARA.ev_tol.offset, GCGE_PARA.orth_para.offset, GCGE_PARA.cg_rate.offset,
            GCGE_PARA.res.offset, GCGE_PARA.stat_para.offset, GCGE_PARA.opt_allgatherv.offset,
            GCGE_PARA.stagnation_ratio.offset, C.sizeof(GCGE_ORTH_PARA), C.sizeof(GCGE_WORKSPACE),
            GCGE_WORKSPACE.unlock.offset, GCGE_WORKSPACE.eval.offset, C.sizeof(GCGE_SOLVER),
            C.sizeof(GCGE_OPS), GCGE_OPS.ArrayScale.offset, GCGE_OPS.MultiVecAxpbyDiag.offset,
            GCGE_PARA.lock_converged.offset, C.sizeof(GCGE_HIP_MIXED_OPTS),
            GCGE_HIP_MIXED_OPTS.stagnation_start.offset, C.sizeof(GCGE_HIP_MIXED_INFO)]
    return list(out[:n]), mine
