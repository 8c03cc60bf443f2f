"""The drop-in boundary itself.

1. The reference's own solver core (src/*.c, built from its sources into
   oracle/_ref/libgcge_ref.so, here only as the checker) drives the HIP
   backend through the 35 reference slots of GCGE_OPS (src/gcge_ops.h:29-215)
   filled by GCGE_HIP_SetOps: GCGE_SOLVER_Create/Setup/Solve of the
   reference, matrices and multivectors of app/hip.  It must reach the
   eigenvalues of the reference's own app/csr run (golden fixtures) within
   1e-10 relative, in the same number of outer iterations (within 2).
2. A user LinearSolver slot (src/gcge_xpw.c:177-188) replaces the block CG of
   the W step: it is called once per W column and the solve converges.
Both libraries are loaded RTLD_LOCAL, so each resolves its own GCGE_*
symbols.
"""
import ctypes as C
import os

import numpy as np
import pytest

from gcge_hip import Matrix, capi, solve
from gcge_hip import problems as P

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libgcge_ref.so")
GOLD = os.path.join(ROOT, "tests", "golden")
N_REF_SLOTS = 36  # VecSetRandomValue .. ArrayScale, linear_solver_workspace included


def _ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libgcge_ref.so not built (make -C oracle ref)")
    return C.CDLL(REF_LIB, mode=os.RTLD_LOCAL)


class RefSolver(C.Structure):
    # src/gcge_solver.h:28-39
    _fields_ = [("A", C.c_void_p), ("B", C.c_void_p), ("evec", C.c_void_p), ("eval", capi.DP),
                ("para", C.c_void_p), ("ops", C.c_void_p), ("workspace", C.c_void_p)]


def _load(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    import scipy.sparse as sp
    n = int(d["n"])
    A = sp.csr_matrix((d["A_data"], d["A_indices"], d["A_indptr"]), shape=(n, n))
    B = sp.csr_matrix((d["B_data"], d["B_indices"], d["B_indptr"]), shape=(n, n)) if bool(d["has_B"]) else None
    return d, A, B


@pytest.mark.parametrize("name,flags", [
    ("lap2d32_x0", ["-gcge_ev_tol", "1e-10", "-gcge_ev_max_it", "80"]),
    ("lap3dmass8_x0", ["-gcge_ev_max_it", "60"]),
    ("vardiff2d16_x0", ["-gcge_ev_max_it", "60"]),
])
def test_reference_core_drives_hip_ops(name, flags):
    ref = _ref()
    hip = capi.load()
    d, A, B = _load(name)
    n, nev = int(d["n"]), int(d["nev"])
    mA = Matrix(A)
    mB = Matrix(B) if B is not None else None
    # the reference's solver object, its GCGE_OPS prefix filled by app/hip
    sp_ = C.POINTER(RefSolver)()
    ref.GCGE_SOLVER_Create(C.byref(sp_))
    s = sp_.contents
    ours = capi.GCGE_OPS()
    hip.GCGE_HIP_SetOps(C.byref(ours))
    C.memmove(s.ops, C.addressof(ours), N_REF_SLOTS * C.sizeof(C.c_void_p))
    ref.GCGE_SOLVER_SetMatA(sp_, C.c_void_p(mA.ptr))
    if mB is not None:
        ref.GCGE_SOLVER_SetMatB(sp_, C.c_void_p(mB.ptr))
    ref.GCGE_SOLVER_SetNumEigen(sp_, nev)
    args = [b"gcge", b"-gcge_given_init_evec", b"1", b"-gcge_print_para", b"0", b"-gcge_print_result", b"0"] + \
        [f.encode() for f in flags]
    argv = (C.c_char_p * len(args))(*args)
    ref.GCGE_PARA_SetFromCommandLine(C.c_void_p(s.para), len(args), argv)
    evec = C.c_void_p()
    hip.HIP_MultiVecCreate(C.byref(evec), nev, mA.ptr)
    X0 = np.asfortranarray(d["X0"])
    hip.HIP_MultiVecSetFromHost(evec, 0, nev, X0.ctypes.data_as(capi.DP), n)
    capi.raise_on_error("initial block")
    ev = np.zeros(nev)
    ref.GCGE_SOLVER_SetEigenvalues(sp_, ev.ctypes.data_as(capi.DP))
    ref.GCGE_SOLVER_SetEigenvectors(sp_, evec)
    ref.GCGE_SOLVER_Setup(sp_)
    ref.GCGE_SOLVER_Solve(sp_)
    hip.GCGE_HIP_Synchronize()
    capi.raise_on_error("reference core over app/hip")
    # GCGE_PARA of the reference: num_iter / num_unlock read through the same layout
    para = C.cast(C.c_void_p(s.para), C.POINTER(capi.GCGE_PARA)).contents
    X = np.zeros((n, nev), order="F")
    hip.HIP_MultiVecGetToHost(evec, 0, nev, X.ctypes.data_as(capi.DP), n)
    np.testing.assert_allclose(ev, d["eval"], rtol=1e-10)
    assert para.num_unlock == int(d["num_unlock"]) == 0
    assert abs(para.num_iter - int(d["num_iter"])) <= 2, (para.num_iter, int(d["num_iter"]))
    BX = X if B is None else B @ X
    R = A @ X - BX * ev[None, :]
    res = np.linalg.norm(R, axis=0) / np.linalg.norm(X, axis=0) / np.maximum(1.0, np.abs(ev))
    tol = float(flags[1]) if flags[0] == "-gcge_ev_tol" else 1e-8
    assert res.max() <= tol * 1.5
    ref.GCGE_SOLVER_Free(C.byref(sp_))   # not Free_All: eval is this test's array
    hip.HIP_MultiVecDestroy(C.byref(evec))
    mA.close()
    if mB is not None:
        mB.close()


LinearSolverFn = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)


def test_user_linear_solver_slot():
    """A LinearSolver in the ops table takes the W solves: here 25 steps of
    plain CG written with the per-vector slots (MatDotVec, VecInnerProd,
    VecAxpby, VecCreateByVec).  It is called once per W column, and the
    outer iteration reaches the closed-form eigenvalues."""
    from gcge_hip import api
    hip = capi.load()
    A, _ = P.lap2d(16)
    nev = 6
    mA = Matrix(A)
    calls = [0]
    holder = {}

    def linear_solver(mat, b, x, ops_p):
        calls[0] += 1
        o = C.cast(ops_p, C.POINTER(capi.GCGE_OPS)).contents
        VecAxpby = C.CFUNCTYPE(None, C.c_double, C.c_void_p, C.c_double, C.c_void_p)(o.VecAxpby)
        VecIP = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, capi.DP)(o.VecInnerProd)
        MatDotVec = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)(o.MatDotVec)
        VecCreate = C.CFUNCTYPE(None, C.POINTER(C.c_void_p), C.c_void_p)(o.VecCreateByVec)
        VecDestroy = C.CFUNCTYPE(None, C.POINTER(C.c_void_p))(o.VecDestroy)
        r, p, w = C.c_void_p(), C.c_void_p(), C.c_void_p()
        for v in (r, p, w):
            VecCreate(C.byref(v), x)
        MatDotVec(mat, x, r)
        VecAxpby(1.0, b, -1.0, r)          # r = b - A x
        VecAxpby(1.0, r, 0.0, p)
        rr = C.c_double()
        VecIP(r, r, C.byref(rr))
        for _ in range(25):
            if rr.value <= 0.0:
                break
            MatDotVec(mat, p, w)
            pw = C.c_double()
            VecIP(p, w, C.byref(pw))
            alpha = rr.value / pw.value
            VecAxpby(alpha, p, 1.0, x)
            VecAxpby(-alpha, w, 1.0, r)
            rr_new = C.c_double()
            VecIP(r, r, C.byref(rr_new))
            VecAxpby(1.0, r, rr_new.value / rr.value, p)
            rr = rr_new
        for v in (r, p, w):
            VecDestroy(C.byref(v))

    cb = LinearSolverFn(linear_solver)
    holder["cb"] = cb
    S = api.Solver(mA, None, nev, precision="fp64", ev_tol=1e-9, conv_type="NS", ev_max_it=200)
    S.solver.contents.ops.contents.LinearSolver = C.cast(cb, C.c_void_p)
    r = S.solve(P.random_x0(A.shape[0], nev, 3))
    S.close()
    mA.close()
    assert r["num_unlock"] == 0
    np.testing.assert_allclose(r["eval"], P.lap2d_eigs(16, nev), rtol=1e-10)
    assert calls[0] >= nev  # at least one W block solved column by column
    assert r["bcg_calls"] == 0  # the block CG did not run
