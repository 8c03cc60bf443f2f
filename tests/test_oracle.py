"""Pins the oracle (C restatement of the reference's solve path).

1. Bit-exact agreement with the reference's own outputs (golden fixtures made
   by tests/golden/make_golden.py from the reference built from its sources):
   eigenvalues, residuals, eigenvectors, iteration counts and the whole
   per-iteration eigenvalue/residual history.
2. Closed-form spectra (2-D Laplacian as in the reference's
   test/hypre/laplace.c:31-43; 3-D Laplacian with P1 mass).
3. Dense LAPACK (scipy.linalg.eigh) on the reference's test matrices.
4. Invariants: B-orthonormal eigenvectors, residuals below tolerance.
5. When oracle/_ref is built here, live runs of the reference on fresh seeds.
"""
import ast
import glob
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import oracle_api as O
from gcge_hip import problems as P

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load_case(path):
    d = np.load(path)
    n = int(d["n"])
    A = sp.csr_matrix((d["A_data"], d["A_indices"], d["A_indptr"]), shape=(n, n))
    B = None
    if bool(d["has_B"]):
        B = sp.csr_matrix((d["B_data"], d["B_indices"], d["B_indptr"]), shape=(n, n))
    X0 = d["X0"] if bool(d["has_X0"]) else None
    kw = ast.literal_eval(str(d["orc_kwargs"]))
    return d, A, B, X0, kw


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_oracle_matches_reference_bit_exact(path, tmp_path):
    d, A, B, X0, kw = load_case(path)
    tr = str(tmp_path / "trace.txt")
    r = O.solve(A, B, int(d["nev"]), X0=X0, trace_path=tr, seed=int(d["seed"]), **kw)
    assert r["info"]["num_iter"] == int(d["num_iter"])
    assert r["info"]["num_unlock"] == int(d["num_unlock"])
    np.testing.assert_array_equal(r["eval"], d["eval"])
    np.testing.assert_array_equal(r["res"], d["res"])
    np.testing.assert_array_equal(r["evec"], d["evec"])
    its, notices = O.parse_trace(open(tr).read())
    assert notices == int(d["notices"])
    np.testing.assert_array_equal(np.array([i["num_unlock"] for i in its]), d["hist_unlock"])
    if len(its):
        # the history is printed with 16 significant digits (%20.15e)
        np.testing.assert_array_equal(np.array([i["eval"] for i in its]), d["hist_eval"])
        np.testing.assert_array_equal(np.array([i["res"] for i in its]), d["hist_res"])


@pytest.mark.parametrize("nx,nev", [(16, 6), (32, 20)])
def test_oracle_closed_form_lap2d(nx, nev):
    A, _ = P.lap2d(nx)
    X0 = P.random_x0(A.shape[0], nev, 1)
    r = O.solve(A, None, nev, X0=X0, ev_tol=1e-11, ev_max_it=200)
    ex = P.lap2d_eigs(nx, nev)
    np.testing.assert_allclose(r["eval"], ex, rtol=1e-10)


def test_oracle_closed_form_lap3d_mass():
    A, B = P.lap3d_mass(8)
    X0 = P.random_x0(A.shape[0], 10, 2)
    r = O.solve(A, B, 10, X0=X0, ev_tol=1e-12, ev_max_it=200)
    np.testing.assert_allclose(r["eval"], P.lap3d_mass_eigs(8, 10), rtol=1e-10)


def test_closed_form_helpers_against_dense():
    """The closed forms themselves are checked against dense LAPACK."""
    A, _ = P.lap2d(12)
    np.testing.assert_allclose(P.lap2d_eigs(12, 15), sla.eigh(A.toarray(), eigvals_only=True)[:15],
                               rtol=1e-12)
    A, B = P.lap3d_mass(6)
    ev = sla.eigh(A.toarray(), B.toarray(), eigvals_only=True)[:12]
    np.testing.assert_allclose(P.lap3d_mass_eigs(6, 12), ev, rtol=1e-11)


def _dense_ref(A, B, nev):
    """Lowest nev eigenvalues by dense LAPACK.  Where B has zero rows (the
    Dirichlet rows of the reference's FEM test matrices) the finite spectrum
    is that of the block on the remaining rows."""
    if B is None:
        return sla.eigh(A.toarray(), eigvals_only=True)[:nev]
    rb = np.asarray(abs(B).sum(axis=1)).ravel()
    I = np.where(rb > 1e-20)[0]
    Ad = A.toarray()[np.ix_(I, I)]
    Bd = B.toarray()[np.ix_(I, I)]
    return sla.eigh(Ad, Bd, eigvals_only=True)[:nev]


@pytest.mark.parametrize("name", ["A3M3_rand", "vardiff2d16_x0", "lap3dmass8_x0", "testA_nev1"])
def test_oracle_against_dense_lapack(name):
    d, A, B, X0, kw = load_case(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    nev = int(d["nev"])
    r = O.solve(A, B, nev, X0=X0, seed=int(d["seed"]), **kw)
    np.testing.assert_allclose(r["eval"], _dense_ref(A, B, nev), rtol=1e-8)


def test_oracle_dirichlet_case_against_dense_interior():
    """A_5/M_5 with Dirichlet rows: M_5 has zero rows there, and the finite
    spectrum is that of the interior block."""
    d, A, B, X0, kw = load_case(os.path.join(os.path.dirname(__file__), "golden",
                                             "A5M5_example.npz"))
    rb = np.asarray(abs(B).sum(axis=1)).ravel()
    I = np.where(rb > 1e-20)[0]
    Ad = A.toarray()[np.ix_(I, I)]
    Bd = B.toarray()[np.ix_(I, I)]
    ex = sla.eigh(Ad, Bd, eigvals_only=True)[:10]
    r = O.solve(A, B, 10, seed=int(d["seed"]), dirichlet_boundary=1, ev_tol=1e-11, ev_max_it=100)
    np.testing.assert_allclose(r["eval"], ex, rtol=1e-9)


@pytest.mark.parametrize("name", ["lap3dmass8_x0", "vardiff2d16_x0", "A3M3_rand"])
def test_oracle_invariants(name):
    d, A, B, X0, kw = load_case(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    nev = int(d["nev"])
    r = O.solve(A, B, nev, X0=X0, seed=int(d["seed"]), **kw)
    X = r["evec"]
    BX = X if B is None else B @ X
    np.testing.assert_allclose(X.T @ BX, np.eye(nev), atol=1e-10)
    R = A @ X - BX * r["eval"][None, :]
    rel = np.linalg.norm(R, axis=0) / np.linalg.norm(X, axis=0) / np.maximum(1.0, abs(r["eval"]))
    assert np.all(rel <= 1e-8 * 1.0001)
    # Courant-Fischer: Ritz values are upper bounds of the exact ones
    assert np.all(r["eval"] >= _dense_ref(A, B, nev) * (1 - 1e-12))


def test_oracle_ns_criterion_extension():
    """conv_type NS (extension): ||Ax - lam Bx|| / (|lam| ||Bx||) <= tol on exit."""
    A, B = P.lap3d_mass(6)
    X0 = P.random_x0(A.shape[0], 8, 3)
    r = O.solve(A, B, 8, X0=X0, conv_type="NS", ev_tol=1e-9, ev_max_it=200)
    X = r["evec"]
    BX = B @ X
    R = A @ X - BX * r["eval"][None, :]
    ns = np.linalg.norm(R, axis=0) / (abs(r["eval"]) * np.linalg.norm(BX, axis=0))
    # residuals of pairs that converged earlier are not recomputed
    # (gcge_eigsol.c:574-652 loops over unlock[] only); all are below tol
    assert np.all(r["res"] <= 1e-9)
    assert np.all(ns <= 1e-9)


def test_kernel_restatements():
    rng = np.random.default_rng(0)
    A, _ = P.vardiff2d(20, seed=1)
    n = A.shape[0]
    X = rng.standard_normal((n, 5))
    np.testing.assert_allclose(O.csr_matmultivec(A, X), A @ X, rtol=0, atol=1e-13)
    V = rng.standard_normal((n, 7))
    np.testing.assert_allclose(O.multivec_innerprod(V, X), V.T @ X, rtol=1e-12, atol=1e-12)
    c = rng.standard_normal((7, 3))
    np.testing.assert_allclose(O.multivec_lincomb(V, c), V @ c, rtol=1e-12, atol=1e-12)


def test_bcg_restatement_reduces_residual():
    A, _ = P.lap2d(20)
    n = A.shape[0]
    rng = np.random.default_rng(1)
    Xt = rng.standard_normal((n, 4))
    RHS = A @ Xt
    W, R, it = O.bcg(A, RHS, np.zeros((n, 4)), cg_max_it=200, cg_rate=1e-10)
    np.testing.assert_allclose(W, Xt, atol=1e-7)
    # final residual block equals RHS - A W
    np.testing.assert_allclose(R, RHS - A @ W, atol=1e-8)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs /root/reference)")
@pytest.mark.parametrize("seed", [21, 22])
def test_oracle_matches_live_reference(seed, tmp_path):
    A, B = P.lap3d_mass(7)
    X0 = P.random_x0(A.shape[0], 9, seed)
    r = O.ref_solve(A, B, 9, X0=X0, flags=["-gcge_ev_max_it", "40"])
    o = O.solve(A, B, 9, X0=X0, ev_max_it=40)
    np.testing.assert_array_equal(o["eval"], r["eval"])
    np.testing.assert_array_equal(o["evec"], r["evec"])
    assert o["info"]["num_iter"] == r["info"]["num_iter"]
