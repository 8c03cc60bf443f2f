"""Matrices uploaded in a locality row order (HIP_MatCreateFromCSROrdered):
the device holds P A P^T and every multivector created from it holds rows
in that order, while every host transfer, the random initial vectors, the
value updates and the gradient Jacobians keep the caller's order.  Products
are bit-identical row by row to the caller's matrix; solves reach the same
eigenpairs (1e-10) as with the natural order; the reference's fixtures are
met in the same number of outer iterations (within 2)."""
import os

import numpy as np
import pytest

import oracle_api as O
from gcge_hip import Matrix, MultiVec, Ops, capi, eig_gradient_generalized, locality_order, solve
from gcge_hip import problems as P

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_ordered_spmm_and_transfers_in_callers_order():
    A, B = P.lap3d_mass(12)
    n = A.shape[0]
    mA = Matrix(A, order="locality")
    perm = mA.order()
    assert perm is not None and np.array_equal(np.sort(perm), np.arange(n))
    mB = Matrix(B, order=perm)
    X = np.random.default_rng(1).standard_normal((n, 7))
    ops = Ops()
    for M, mM in ((A, mA), (B, mB)):
        for prec in (capi.GCGE_HIP_FP64, capi.GCGE_HIP_FP32):
            Xp = X if prec == capi.GCGE_HIP_FP64 else X.astype(np.float32).astype(np.float64)
            x, y = MultiVec(mM, 7, prec, data=Xp), MultiVec(mM, 7, prec)
            np.testing.assert_array_equal(x.get(), Xp)          # round trip in the caller's order
            ops.matdot(mM, x, y)
            ref = O.csr_matmultivec(M, Xp)
            if prec == capi.GCGE_HIP_FP64:
                np.testing.assert_array_equal(y.get(), ref)      # row sums in the same order
            else:
                np.testing.assert_allclose(y.get(), ref, rtol=2e-5, atol=2e-5 * np.abs(ref).max())
    ops.close()
    mA.close()
    mB.close()


@pytest.mark.parametrize("order", ["locality", "random"])
@pytest.mark.parametrize("name", ["lap2d32_x0", "lap3dmass8_x0", "vardiff2d16_x0"])
def test_ordered_solve_matches_reference_fixture(name, order):
    """The reference's fixtures solved on P A P^T: the locality order (16 x 16
    tiles of the 32 x 32 grid; the identity for grids of one tile) and a
    random permutation of the rows."""
    import ast
    import scipy.sparse as sp
    d = np.load(os.path.join(GOLD, name + ".npz"))
    n = int(d["n"])
    A = sp.csr_matrix((d["A_data"], d["A_indices"], d["A_indptr"]), shape=(n, n))
    B = sp.csr_matrix((d["B_data"], d["B_indices"], d["B_indptr"]), shape=(n, n)) if bool(d["has_B"]) else None
    kw = ast.literal_eval(str(d["orc_kwargs"]))
    mA = Matrix(A, order="locality" if order == "locality" else np.random.default_rng(3).permutation(n))
    mB = Matrix(B, order=mA.order()) if B is not None else None
    r = solve(mA, mB, int(d["nev"]), X0=d["X0"], **kw)
    np.testing.assert_allclose(r["eval"], d["eval"], rtol=1e-10)
    assert r["num_unlock"] == 0
    assert abs(r["num_iter"] - int(d["num_iter"])) <= 2
    # eigenvectors come back in the caller's order
    BX = r["evec"] if B is None else B @ r["evec"]
    R = A @ r["evec"] - BX * r["eval"][None, :]
    res = np.linalg.norm(R, axis=0) / np.linalg.norm(r["evec"], axis=0) / np.maximum(1.0, np.abs(r["eval"]))
    assert res.max() <= kw.get("ev_tol", 1e-8) * 1.5
    mA.close()
    if mB is not None:
        mB.close()


def test_ordered_random_start_and_mixed_solve():
    """The rand() initial block is drawn in the caller's order, so an ordered
    and a natural matrix start from the same vectors; the mixed path runs on
    ordered matrices too."""
    A, _ = P.vardiff3d(14, seed=2)
    nev = 10
    kw = dict(conv_type="NS", ev_tol=1e-9, ev_max_it=300)
    r0 = solve(Matrix(A), None, nev, seed=5, **kw)
    r1 = solve(Matrix(A, order="locality"), None, nev, seed=5, **kw)
    np.testing.assert_allclose(r1["eval"], r0["eval"], rtol=1e-10)
    assert abs(r1["num_iter"] - r0["num_iter"]) <= 2
    X0 = P.random_x0(A.shape[0], nev, 6)
    rm = solve(Matrix(A, order="locality"), None, nev, X0=X0, precision="mixed", **kw)
    assert rm["num_unlock"] == 0 and rm["rc"] == 0
    np.testing.assert_allclose(rm["eval"], r0["eval"], rtol=1e-10)


def test_ordered_values_and_gradient_in_callers_entry_order():
    m = 24
    kappa = P.vardiff2d_theta(m, var=0.25, seed=8)
    A, _ = P.vardiff2d(m, theta=kappa, anis=1.3)
    nev = 6
    kw = dict(conv_type="NS", ev_tol=1e-11, ev_max_it=500)
    X0 = P.random_x0(A.shape[0], nev, 9)
    mN, mO = Matrix(A), Matrix(A, order="locality")
    rN = solve(mN, None, nev, X0=X0, **kw)
    rO = solve(mO, None, nev, X0=X0, **kw)
    np.testing.assert_allclose(rO["eval"], rN["eval"], rtol=1e-10)
    ja = P.fv_diffusion_jacobian(kappa, A, face_scale=(1.0, 1.3))
    gN = eig_gradient_generalized(mN, None, rN["evec"], rN["eval"], np.ones(nev), m * m, ja)
    gO = eig_gradient_generalized(mO, None, rO["evec"], rO["eval"], np.ones(nev), m * m, ja)
    np.testing.assert_allclose(gO, gN, rtol=1e-7, atol=1e-12 * np.abs(gN).max())
    # new values in the caller's entry order
    mO.set_values(3.0 * A.data)
    rO3 = solve(mO, None, nev, X0=X0, **kw)
    np.testing.assert_allclose(rO3["eval"], 3.0 * rN["eval"], rtol=1e-10)
    mN.close()
    mO.close()


def test_mismatched_orders_are_refused():
    A, B = P.lap3d_mass(8)
    mA = Matrix(A, order="locality")
    mB = Matrix(B)
    with pytest.raises(RuntimeError, match="171"):
        solve(mA, mB, 4, X0=P.random_x0(A.shape[0], 4, 1))
    mA.close()
    mB.close()
