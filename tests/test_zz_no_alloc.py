"""The no-allocation entry point GCGE_HIP_MultiVecInnerProdDiagNoAlloc
(include/gcge_app_hip.h) against the MultiVecInnerProdDiag slot it varies:
the same kernels and summation order, so identical sums, on two calls in a
row (the second smaller, reusing the file-scope scratch array), in FP64 and
FP32, plus its argument checks."""
import ctypes as C

import numpy as np
import pytest

from gcge_hip import Matrix, MultiVec, Ops, capi
from gcge_hip import problems as P

pytestmark = pytest.mark.gpu


def _noalloc():
    L = capi.load()
    f = L.GCGE_HIP_MultiVecInnerProdDiagNoAlloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, capi.DP, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    return f


def _call(f, X, Y, xs, ys, k):
    out = np.full(k, np.nan)
    s = (C.c_int * 2)(xs, ys)
    e = (C.c_int * 2)(xs + k, ys + k)
    rc = f(X.ptr, Y.ptr, out.ctypes.data_as(capi.DP), s, e)
    return rc, out


@pytest.mark.parametrize("prec", [capi.GCGE_HIP_FP64, capi.GCGE_HIP_FP32])
def test_noalloc_dots_match_slot_two_calls(prec):
    A, _ = P.vardiff2d(300, seed=1)        # n = 90,000: 256 partial blocks per column
    n = A.shape[0]
    mA = Matrix(A)
    rng = np.random.default_rng(3)
    X = MultiVec(mA, 128, prec, data=rng.standard_normal((n, 128)))
    Y = MultiVec(mA, 130, prec, data=rng.standard_normal((n, 130)))
    ops = Ops()
    f = _noalloc()
    for xs, ys, k in ((0, 2, 128), (5, 0, 7)):   # second call smaller
        rc, got = _call(f, X, Y, xs, ys, k)
        assert rc == 0, capi.last_error()
        s = (C.c_int * 2)(xs, ys)
        e = (C.c_int * 2)(xs + k, ys + k)
        ref = np.zeros(k)
        ops._f("MultiVecInnerProdDiag", capi.F_MultiVecInnerProdDiag)(X.ptr, Y.ptr, ref.ctypes.data_as(capi.DP), s,
                                                                       e, ops._p)
        capi.raise_on_error("slot")
        np.testing.assert_array_equal(got, ref)
        xh = X.get(xs, k)
        yh = Y.get(ys, k)
        np.testing.assert_allclose(got, np.einsum("ij,ij->j", xh, yh), rtol=1e-12 if prec == 0 else 1e-6)
    ops.close()
    X.close()
    Y.close()
    mA.close()


def test_noalloc_rejects_too_many_columns_and_bad_ranges():
    A, _ = P.lap2d(16)
    mA = Matrix(A)
    X = MultiVec(mA, 140)
    f = _noalloc()
    rc, _ = _call(f, X, X, 0, 0, 129)
    assert rc == 202
    code, msg = capi.last_error()
    assert code == 202 and "limit of 128" in msg
    capi.load().GCGE_HIP_ClearError()
    rc, _ = _call(f, X, X, 100, 0, 41)
    assert rc == 201
    capi.load().GCGE_HIP_ClearError()
    X.close()
    mA.close()
