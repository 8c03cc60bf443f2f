"""Kernel-level parity of the HIP backend, through the GCGE_OPS slots.

Element-wise kernels (SpMM, y = a x + b y, the CG update, copies) must be
bit-identical to the CPU restatement of the reference's loops; reductions
(dots, Gram, linear combinations, residual norms) are checked against the
oracle within a stated relative tolerance, since their summation order
differs by design.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_api as O
from gcge_hip import Matrix, MultiVec, Ops, capi, eig_gradient
from gcge_hip import problems as P

pytestmark = pytest.mark.gpu

F64, F32 = capi.GCGE_HIP_FP64, capi.GCGE_HIP_FP32
RTOL_RED = 1e-12      # FP64 reductions vs the oracle's sequential sums
RTOL_F32 = 2e-5       # FP32 kernels vs FP64


@pytest.fixture(scope="module")
def ops():
    o = Ops()
    yield o
    o.close()


def rng(s=0):
    return np.random.default_rng(s)


def mats():
    A2, _ = P.vardiff2d(40, seed=2)                  # 5-point, 1600 rows
    A3, B3 = P.lap3d_mass(11)                       # 7-point and 27-point (global path)
    G, _ = P.chung_lu(3000, mean_deg=8.0, seed=1)    # power-law rows
    return {"vardiff2d": A2, "lap3d": A3, "mass27": B3, "chunglu": G}


MATS = mats()


def _matrix_with_gather(A, layout, monkeypatch):
    # the SpMM gather layout is chosen when the matrix is created
    monkeypatch.setenv("GCGE_HIP_SPMM_GATHER", layout)
    m = Matrix(A)
    monkeypatch.delenv("GCGE_HIP_SPMM_GATHER")
    return m


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("k", [1, 3, 4, 70])
def test_spmm_fp64_bit_exact(ops, monkeypatch, name, k):
    # column-gather path (the row-gather path has its own tests below)
    A = MATS[name]
    n = A.shape[0]
    X = rng(k).standard_normal((n, k))
    mA = _matrix_with_gather(A, "columns", monkeypatch)
    x = MultiVec(mA, k, data=X)
    y = MultiVec(mA, k)
    ops.matdot(mA, x, y)
    np.testing.assert_array_equal(y.get(), O.csr_matmultivec(A, X))


@pytest.mark.parametrize("name", ["vardiff2d", "mass27", "chunglu"])
@pytest.mark.parametrize("layout", ["columns", "rows"])
def test_spmm_dot_fused(ops, monkeypatch, name, layout):
    A = MATS[name]
    n, k = A.shape[0], 6
    X = rng(3).standard_normal((n, k))
    mA = _matrix_with_gather(A, layout, monkeypatch)
    x = MultiVec(mA, k, data=X)
    y = MultiVec(mA, k)
    d = ops.matdot_dot(mA, x, y)
    Y = O.csr_matmultivec(A, X)
    np.testing.assert_array_equal(y.get(), Y)
    np.testing.assert_allclose(d, np.einsum("ij,ij->j", X, Y), rtol=RTOL_RED)


def test_spmm_fp32(ops):
    A = MATS["lap3d"]
    n, k = A.shape[0], 5
    X = rng(4).standard_normal((n, k)).astype(np.float32).astype(np.float64)
    mA = Matrix(A)
    x = MultiVec(mA, k, F32, data=X)
    y = MultiVec(mA, k, F32)
    ops.matdot(mA, x, y)
    ref = A @ X
    np.testing.assert_allclose(y.get(), ref, rtol=RTOL_F32, atol=RTOL_F32 * np.abs(ref).max())


def test_spmm_column_offsets(ops):
    A = MATS["vardiff2d"]
    n = A.shape[0]
    X = rng(5).standard_normal((n, 9))
    mA = Matrix(A)
    x = MultiVec(mA, 9, data=X)
    y = MultiVec(mA, 12)
    ops.matdot(mA, x, y, xs=2, ys=5, k=6)
    Y = y.get()
    np.testing.assert_array_equal(Y[:, 5:11], O.csr_matmultivec(A, X[:, 2:8]))
    assert np.all(Y[:, :5] == 0) and np.all(Y[:, 11:] == 0)


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("k", [2, 3, 5, 8, 16, 17, 33, 64, 70, 128, 130])
def test_spmm_row_gather_fp64_bit_exact(ops, monkeypatch, name, k):
    # row-major gather path (staged X rows, lanes over columns, long rows one
    # wavefront each): same CSR-order sums as the column path, so the same bits
    A = MATS[name]
    n = A.shape[0]
    X = rng(k + 11).standard_normal((n, k))
    mA = _matrix_with_gather(A, "rows", monkeypatch)
    x = MultiVec(mA, k, data=X)
    y = MultiVec(mA, k)
    ops.matdot(mA, x, y)
    Y = O.csr_matmultivec(A, X)
    np.testing.assert_array_equal(y.get(), Y)
    y2 = MultiVec(mA, k)
    d = ops.matdot_dot(mA, x, y2)
    np.testing.assert_array_equal(y2.get(), Y)
    np.testing.assert_allclose(d, np.einsum("ij,ij->j", X, Y), rtol=RTOL_RED)


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("kc", ["", "1", "2", "4", "8"])
def test_spmm_halo_fp64_bit_exact(ops, monkeypatch, name, kc):
    """Staged row blocks (k_spmm_halo): every X row a 256-row block reads is
    copied to LDS once per column pass; sums in CSR order from 0, so the same
    bits as the oracle's CSR_MatDotVec loop.  mass27 exercises rows longer
    than the register-held entries, chunglu the long-row kernel beside it;
    kc forces the columns per LDS pass."""
    A = MATS[name]
    n = A.shape[0]
    if kc:
        monkeypatch.setenv("GCGE_HIP_HALO_KC", kc)
    mA = _matrix_with_gather(A, "halo", monkeypatch)
    for k in (1, 3, 9, 70, 130):
        X = rng(k + 21).standard_normal((n, k))
        x = MultiVec(mA, k, data=X)
        y, y2 = MultiVec(mA, k), MultiVec(mA, k)
        ops.matdot(mA, x, y)
        Y = O.csr_matmultivec(A, X)
        np.testing.assert_array_equal(y.get(), Y)
        d = ops.matdot_dot(mA, x, y2)
        np.testing.assert_array_equal(y2.get(), Y)
        np.testing.assert_allclose(d, np.einsum("ij,ij->j", X, Y), rtol=RTOL_RED)
        # FP32 storage: identical to the direct-gather kernel's
        xf = MultiVec(mA, k, F32, data=X)
        yf = MultiVec(mA, k, F32)
        ops.matdot(mA, xf, yf)
        mC = _matrix_with_gather(A, "columns", monkeypatch)
        xc = MultiVec(mC, k, F32, data=X)
        yc = MultiVec(mC, k, F32)
        ops.matdot(mC, xc, yc)
        np.testing.assert_array_equal(yf.get(), yc.get())
        for m in (x, y, y2, xf, yf, xc, yc):
            m.close()
        mC.close()
    # column offsets
    X = rng(14).standard_normal((n, 9))
    x, y = MultiVec(mA, 9, data=X), MultiVec(mA, 12)
    ops.matdot(mA, x, y, xs=2, ys=5, k=6)
    Y = y.get()
    np.testing.assert_array_equal(Y[:, 5:11], O.csr_matmultivec(A, X[:, 2:8]))
    assert np.all(Y[:, :5] == 0) and np.all(Y[:, 11:] == 0)
    mA.close()


@pytest.mark.parametrize("name", ["vardiff2d", "chunglu"])
@pytest.mark.parametrize("k", [2, 6, 40, 100])
def test_spmm_row_gather_fp32_matches_column_gather(ops, monkeypatch, name, k):
    A = MATS[name]
    n = A.shape[0]
    X = rng(k + 12).standard_normal((n, k))
    out = []
    for layout in ("rows", "columns"):
        mA = _matrix_with_gather(A, layout, monkeypatch)
        x = MultiVec(mA, k, F32, data=X)
        y, y2 = MultiVec(mA, k, F32), MultiVec(mA, k, F32)
        ops.matdot(mA, x, y)
        d = ops.matdot_dot(mA, x, y2)
        np.testing.assert_array_equal(y.get(), y2.get())
        out.append((y.get(), np.asarray(d)))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_allclose(out[0][1], out[1][1], rtol=1e-12)


def test_spmm_row_gather_column_offsets(ops, monkeypatch):
    A = MATS["chunglu"]
    n = A.shape[0]
    X = rng(13).standard_normal((n, 9))
    mA = _matrix_with_gather(A, "rows", monkeypatch)
    x = MultiVec(mA, 9, data=X)
    y = MultiVec(mA, 12)
    ops.matdot(mA, x, y, xs=2, ys=5, k=6)
    Y = y.get()
    np.testing.assert_array_equal(Y[:, 5:11], O.csr_matmultivec(A, X[:, 2:8]))
    assert np.all(Y[:, :5] == 0) and np.all(Y[:, 11:] == 0)


@pytest.mark.parametrize("a,b", [(0.0, 2.5), (1.0, 0.0), (-3.0, 0.0), (0.7, -1.3), (1.0, 1.0), (0.0, 0.0)])
def test_axpby_branches_bit_exact(ops, a, b):
    A = MATS["vardiff2d"]
    n, k = A.shape[0], 3
    X = rng(6).standard_normal((n, k))
    Y = rng(7).standard_normal((n, k))
    if b == 0.0:
        Y[5, 1] = np.nan        # y is not read when b == 0
    if a == 0.0:
        X[7, 2] = np.inf        # x is not read when a == 0
    mA = Matrix(A)
    x = MultiVec(mA, k, data=X)
    y = MultiVec(mA, k, data=Y)
    ops.axpby(a, x, b, y)
    ref = Y.copy()
    for j in range(k):
        # branch rules of CSR_VecAxpby (gcge_app_csr.c:330-365)
        r = ref[:, j].copy()
        if a == 0.0:
            r = b * r
        elif b == 0.0:
            r = X[:, j].copy() if a == 1.0 else a * X[:, j]
        else:
            r = a * X[:, j] + b * r
        ref[:, j] = r
    np.testing.assert_array_equal(y.get(), ref)


def test_axpby_diag_and_dots(ops):
    A = MATS["vardiff2d"]
    n, k = A.shape[0], 70
    X = rng(8).standard_normal((n, k))
    Y = rng(9).standard_normal((n, k))
    a = rng(10).standard_normal(k)
    b = rng(11).standard_normal(k)
    a[3] = 0.0
    b[4] = 0.0
    a[5], b[5] = 1.0, 0.0
    mA = Matrix(A)
    x = MultiVec(mA, k, data=X)
    y = MultiVec(mA, k, data=Y)
    ops.axpby_diag(a, x, b, y)
    ref = a[None, :] * X + b[None, :] * Y
    ref[:, 3] = b[3] * Y[:, 3]
    ref[:, 4] = a[4] * X[:, 4]
    ref[:, 5] = X[:, 5]
    got = y.get()
    np.testing.assert_array_equal(got, ref)
    d = ops.dots(x, y)
    np.testing.assert_allclose(d, np.einsum("ij,ij->j", X, ref), rtol=RTOL_RED)


@pytest.mark.parametrize("p,q", [(1, 1), (7, 3), (70, 33), (130, 65)])
def test_gram_against_oracle(ops, p, q):
    A = MATS["lap3d"]
    n = A.shape[0]
    V = rng(12).standard_normal((n, p))
    W = rng(13).standard_normal((n, q))
    mA = Matrix(A)
    v = MultiVec(mA, p, data=V)
    w = MultiVec(mA, q, data=W)
    G = ops.gram(v, w)
    ref = O.multivec_innerprod(V, W)
    np.testing.assert_allclose(G, ref, rtol=RTOL_RED, atol=RTOL_RED * np.abs(ref).max())


# thin kernel (p <= 4, q <= 8 and the transpose), the MFMA tile shapes up to
# 32 x 32, the LDS-slab blocks beyond (one or several blocks per side, uneven
# block splits), partial tiles, and a row count that is not a multiple of any
# superstep or slab, on 100,489 rows so the rows split into chunks
BIG = P.lap2d(317)[0]
GRAM_SHAPES = [(1, 1), (1, 8), (8, 1), (4, 4), (3, 7), (7, 3), (6, 6), (12, 13), (16, 16), (17, 5),
               (25, 25), (32, 32), (33, 40), (64, 64), (100, 100), (200, 100), (118, 12), (12, 118), (70, 16),
               (9, 33), (257, 129), (100, 260)]


@pytest.mark.parametrize("p,q", GRAM_SHAPES)
def test_gram_tiles_large_n(ops, p, q):
    n = BIG.shape[0]
    V = rng(30 + p).standard_normal((n, p))
    W = rng(60 + q).standard_normal((n, q))
    mA = Matrix(BIG)
    v = MultiVec(mA, p + 3, data=np.hstack([V, np.ones((n, 3))]))
    w = MultiVec(mA, q + 2, data=np.hstack([np.ones((n, 2)), W]))
    G = ops.gram(v, w, vs=0, ve=p, ws=2, we=2 + q)
    ref = V.T @ W
    np.testing.assert_allclose(G, ref, rtol=RTOL_RED, atol=RTOL_RED * np.sqrt(n) * 10)


@pytest.mark.parametrize("p,q", [(2, 5), (12, 13), (40, 70), (130, 100)])
def test_gram_fp32_tiles(ops, p, q):
    """FP32 data, FP32 accumulation inside a chunk, FP64 across chunks."""
    n = BIG.shape[0]
    V = rng(90 + p).standard_normal((n, p)).astype(np.float32).astype(np.float64)
    W = rng(91 + q).standard_normal((n, q)).astype(np.float32).astype(np.float64)
    mA = Matrix(BIG)
    v = MultiVec(mA, p, F32, data=V)
    w = MultiVec(mA, q, F32, data=W)
    G = ops.gram(v, w)
    bound = 1e-4 * (np.abs(V).T @ np.abs(W))
    assert np.all(np.abs(G - V.T @ W) <= bound)


def test_gram_sym_large_tiles(ops):
    """sym with 64x64 tiles: tiles below the diagonal are not computed."""
    n, q = BIG.shape[0], 150
    V = rng(92).standard_normal((n, q))
    W = V @ rng(93).standard_normal((q, q)) * 1e-2 + V
    mA = Matrix(BIG)
    v = MultiVec(mA, q, data=V)
    w = MultiVec(mA, q, data=W)
    G = ops.gram(v, w, sym=True)
    up = np.triu(V.T @ W)
    ref = up + np.triu(up, 1).T
    np.testing.assert_allclose(G, ref, rtol=RTOL_RED, atol=RTOL_RED * np.sqrt(n) * 10)


def test_gram_sym_mirrors_upper_triangle(ops):
    A = MATS["lap3d"]
    n, q = A.shape[0], 20
    V = rng(14).standard_normal((n, 30))
    W = rng(15).standard_normal((n, q))
    mA = Matrix(A)
    v = MultiVec(mA, 30, data=V)
    w = MultiVec(mA, q, data=W)
    G = ops.gram(v, w, vs=4, ve=24, sym=True)
    full = V[:, 4:24].T @ W
    up = np.triu(full)
    ref = up + np.triu(up, 1).T
    np.testing.assert_allclose(G, ref, rtol=RTOL_RED, atol=1e-12 * np.abs(full).max())


@pytest.mark.parametrize("p,q", [(1, 1), (5, 2), (70, 40), (208, 100)])
def test_lincomb_against_oracle(ops, p, q):
    A = MATS["vardiff2d"]
    n = A.shape[0]
    X = rng(16).standard_normal((n, p))
    Cf = rng(17).standard_normal((p, q))
    mA = Matrix(A)
    x = MultiVec(mA, p, data=X)
    y = MultiVec(mA, q)
    ops.lincomb(x, y, Cf)
    ref = O.multivec_lincomb(X, Cf)
    np.testing.assert_allclose(y.get(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("p,q,alpha,beta", [(1, 17, 1.0, 0.0), (3, 40, -1.0, 1.0), (16, 16, 0.5, 0.0),
                                             (33, 130, -1.0, 1.0), (300, 100, 1.0, 0.0), (2, 3, 2.0, -0.5),
                                             (8, 8, -1.0, 1.0), (5, 7, 1.0, 0.0), (1, 1, 3.0, 0.0),
                                             (9, 8, -1.0, 1.0), (65, 128, -1.0, 1.0)])
def test_lincomb_tiles_large_n(ops, p, q, alpha, beta):
    n = BIG.shape[0]
    X = rng(100 + p).standard_normal((n, p))
    Y = rng(101 + q).standard_normal((n, q))
    Cf = rng(102).standard_normal((p, q))
    mA = Matrix(BIG)
    x = MultiVec(mA, p + 1, data=np.hstack([np.ones((n, 1)), X]))
    y = MultiVec(mA, q + 2, data=np.hstack([Y, np.ones((n, 2))]))
    ops.lincomb(x, y, Cf, xs=1, ys=0, alpha=alpha, beta=beta)
    ref = alpha * (X @ Cf) + (beta * Y if beta != 0.0 else 0.0)
    got = y.get()
    scale = np.abs(X) @ np.abs(Cf)
    assert np.all(np.abs(got[:, :q] - ref) <= 1e-13 * (scale + np.abs(Y)) + 1e-300)
    np.testing.assert_array_equal(got[:, q:], 1.0)


@pytest.mark.parametrize("p,q,beta", [(3, 5, 0.0), (8, 2, 1.0)])
def test_lincomb_fp32_thin(ops, p, q, beta):
    n = BIG.shape[0]
    X = rng(112).standard_normal((n, p)).astype(np.float32).astype(np.float64)
    Y = rng(113).standard_normal((n, q)).astype(np.float32).astype(np.float64)
    Cf = rng(114).standard_normal((p, q))
    mA = Matrix(BIG)
    x = MultiVec(mA, p, F32, data=X)
    y = MultiVec(mA, q, F32, data=Y)
    ops.lincomb(x, y, Cf, alpha=-1.0, beta=beta)
    ref = beta * Y - X @ Cf.astype(np.float32).astype(np.float64)
    assert np.all(np.abs(y.get() - ref) <= 1e-6 * (np.abs(X) @ np.abs(Cf) + np.abs(Y)))


def test_lincomb_fp32_large_n(ops):
    n, p, q = BIG.shape[0], 37, 70
    X = rng(110).standard_normal((n, p)).astype(np.float32).astype(np.float64)
    Cf = rng(111).standard_normal((p, q))
    mA = Matrix(BIG)
    x = MultiVec(mA, p, F32, data=X)
    y = MultiVec(mA, q, F32)
    ops.lincomb(x, y, Cf)
    ref = X @ Cf.astype(np.float32).astype(np.float64)
    assert np.all(np.abs(y.get() - ref) <= 1e-5 * (np.abs(X) @ np.abs(Cf)))


def test_lincomb_axpby_projection(ops):
    """y = y - X C: the fused projection of the subspace orthogonalization."""
    A = MATS["vardiff2d"]
    n, p, q = A.shape[0], 12, 5
    X = rng(18).standard_normal((n, p))
    Y = rng(19).standard_normal((n, q))
    Cf = rng(20).standard_normal((p, q))
    mA = Matrix(A)
    x = MultiVec(mA, p, data=X)
    y = MultiVec(mA, q, data=Y)
    ops.lincomb(x, y, Cf, alpha=-1.0, beta=1.0)
    ref = Y - O.multivec_lincomb(X, Cf)
    np.testing.assert_allclose(y.get(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_lincomb_fp32(ops):
    A = MATS["vardiff2d"]
    n, p, q = A.shape[0], 40, 20
    X = rng(21).standard_normal((n, p))
    Cf = rng(22).standard_normal((p, q))
    mA = Matrix(A)
    x = MultiVec(mA, p, F32, data=X)
    y = MultiVec(mA, q, F32)
    ops.lincomb(x, y, Cf)
    ref = X.astype(np.float32).astype(np.float64) @ Cf
    np.testing.assert_allclose(y.get(), ref, rtol=1e-4, atol=1e-5 * np.abs(ref).max())


def test_cg_update_bit_exact(ops):
    A = MATS["vardiff2d"]
    n, k = A.shape[0], 7
    g = rng(23)
    Xs, Pm, Wm, R = (g.standard_normal((n, k + 3)), g.standard_normal((n, k)), g.standard_normal((n, k)),
                     g.standard_normal((n, k)))
    alpha = g.standard_normal(k)
    alpha[2] = 0.0
    mA = Matrix(A)
    x, p, w, r = (MultiVec(mA, k + 3, data=Xs), MultiVec(mA, k, data=Pm), MultiVec(mA, k, data=Wm),
                  MultiVec(mA, k, data=R))
    rr = ops.cg_update(x, 3, p, w, r, k, alpha)
    xr = Xs.copy()
    rr_ref = R.copy()
    for j in range(k):
        if alpha[j] == 0.0:
            continue
        xr[:, 3 + j] = alpha[j] * Pm[:, j] + 1.0 * xr[:, 3 + j]
        rr_ref[:, j] = -alpha[j] * Wm[:, j] + 1.0 * rr_ref[:, j]
    np.testing.assert_array_equal(x.get(), xr)
    np.testing.assert_array_equal(r.get(), rr_ref)
    np.testing.assert_allclose(rr, np.einsum("ij,ij->j", rr_ref, rr_ref), rtol=RTOL_RED)


def test_cg_split_steps_bit_exact(ops):
    """BlockCGResidualUpdate and BlockCGSearchUpdate (the fused block CG
    steps) against VecAxpby's roundings: r = -a w + r, x = a p + x, then
    p = r + b p for the first kp columns only."""
    A = MATS["vardiff2d"]
    n, kx, kp = A.shape[0], 9, 6
    g = rng(40)
    X, P_, W, R = (g.standard_normal((n, kx)) for _ in range(4))
    al = g.standard_normal(kx)
    be = g.standard_normal(kx)
    be[2] = 0.0
    mA = Matrix(A)
    x, p, w, r = (MultiVec(mA, kx, data=M) for M in (X, P_, W, R))
    rr = ops.cg_residual_update(w, r, kx, al)
    r_ref = (-al)[None, :] * W + R
    np.testing.assert_array_equal(r.get(), r_ref)
    np.testing.assert_allclose(rr, np.einsum("ij,ij->j", r_ref, r_ref), rtol=RTOL_RED)
    ops.cg_search_update(x, 0, p, r, kx, kp, al, be)
    x_ref = al[None, :] * P_ + X
    p_ref = P_.copy()
    p_ref[:, :kp] = r_ref[:, :kp] + be[None, :kp] * P_[:, :kp]
    p_ref[:, 2] = r_ref[:, 2]          # beta == 0: p = r (copy branch)
    np.testing.assert_array_equal(x.get(), x_ref)
    np.testing.assert_array_equal(p.get(), p_ref)
    # the active count may grow between steps (stale ratios of inactive
    # columns): x for j < 3, p for j < 8
    ops.cg_search_update(x, 0, p, r, 3, 8, al, be)
    x2 = x_ref.copy()
    x2[:, :3] = al[None, :3] * p_ref[:, :3] + x_ref[:, :3]
    p2 = p_ref.copy()
    p2[:, :8] = r_ref[:, :8] + be[None, :8] * p_ref[:, :8]
    p2[:, 2] = r_ref[:, 2]
    np.testing.assert_array_equal(x.get(), x2)
    np.testing.assert_array_equal(p.get(), p2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_pcg_steps(ops, prec):
    """The Jacobi-preconditioned block CG steps, z = D^-1 r formed where it
    is used: BlockPCGResidualUpdate leaves r as BlockCGResidualUpdate does
    (bit for bit) and returns r.z and r.r (FP64 accumulation);
    BlockPCGSearchUpdate adds alpha p to x and sets p = z + beta p with the
    same z, 1 / a_ii rounded once in FP64 (then to FP32 for FP32 data).  A
    zero alpha leaves r as it is; a row without a stored diagonal entry is
    not scaled."""
    A = MATS["vardiff2d"].tolil()
    A[5, 5] = 0.0
    A = A.tocsr()
    A.eliminate_zeros()
    n, k, kx = A.shape[0], 7, 4
    g = rng(44)
    W, R, X, P_ = (g.standard_normal((n, k)) for _ in range(4))
    al, be = g.standard_normal(k), g.standard_normal(k)
    be[1] = 0.0
    f32 = prec == "f32"
    T = np.float32 if f32 else np.float64
    pr = F32 if f32 else capi.GCGE_HIP_FP64
    mA = Matrix(A)
    w, r, x, p = (MultiVec(mA, k, pr, data=M) for M in (W, R, X, P_))
    W, R, X, P_ = (M.astype(T) for M in (W, R, X, P_))
    d = A.diagonal()
    dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 1.0).astype(T)
    r_ref = (-al.astype(T))[None, :] * W + R
    z_ref = r_ref * dinv[:, None]
    rz, rr = ops.pcg_residual_update(mA, w, r, k, al)
    np.testing.assert_array_equal(r.get(), r_ref.astype(np.float64))
    rd, zd = r_ref.astype(np.float64), z_ref.astype(np.float64)
    np.testing.assert_allclose(rz, np.einsum("ij,ij->j", rd, zd), rtol=RTOL_RED)
    np.testing.assert_allclose(rr, np.einsum("ij,ij->j", rd, rd), rtol=RTOL_RED)
    rz0, rr0 = ops.pcg_residual_update(mA, w, r, k, np.zeros(k))
    np.testing.assert_array_equal(r.get(), rd)
    np.testing.assert_allclose(rr0, rr, rtol=0)
    np.testing.assert_allclose(rz0, rz, rtol=0)
    ops.pcg_search_update(mA, x, 0, p, r, kx, k, al, be)
    x_ref = X.copy()
    x_ref[:, :kx] = al[None, :kx].astype(T) * P_[:, :kx] + X[:, :kx]
    p_ref = z_ref + be.astype(T)[None, :] * P_
    p_ref[:, 1] = z_ref[:, 1]  # beta == 0: p = z (copy branch)
    np.testing.assert_array_equal(x.get(), x_ref.astype(np.float64))
    np.testing.assert_array_equal(p.get(), p_ref.astype(np.float64))


@pytest.mark.parametrize("withB", [False, True])
def test_residual_norms(ops, withB):
    A, B = P.lap3d_mass(9)
    n, k = A.shape[0], 9
    X = rng(24).standard_normal((n, k))
    lam = rng(25).uniform(1, 100, k)
    idx = np.array([3, 0, 8, 5], dtype=np.int32)
    mA = Matrix(A)
    mB = Matrix(B) if withB else None
    x = MultiVec(mA, k, data=X)
    rn, xn, bn = ops.residual_norms(mA, mB, x, idx, lam[idx])
    BX = (B @ X) if withB else X
    R = A @ X[:, idx] - BX[:, idx] * lam[idx][None, :]
    np.testing.assert_allclose(rn, np.linalg.norm(R, axis=0), rtol=1e-11)
    np.testing.assert_allclose(xn, np.linalg.norm(X[:, idx], axis=0), rtol=1e-12)
    np.testing.assert_allclose(bn, np.linalg.norm(BX[:, idx], axis=0), rtol=1e-12)


@pytest.mark.parametrize("case", ["vardiff3d_k7", "chung_lu_k130", "vardiff3d_f32_k9"])
def test_residual_norms_column_chunks(ops, case):
    """Ragged column counts (not a multiple of the kernel's column chunk, and
    more than one launch's worth), workgroups whose rows hold more nonzeros
    than the LDS stage (power-law graph: read straight from the CSR arrays)
    and FP32 storage, against FP64 numpy."""
    if case.startswith("chung_lu"):
        A, _ = P.chung_lu(6000, seed=3)
    else:
        A, _ = P.vardiff3d(14, seed=2)
    n = A.shape[0]
    k = {"vardiff3d_k7": 7, "chung_lu_k130": 130, "vardiff3d_f32_k9": 9}[case]
    X = rng(40).standard_normal((n, k + 3))
    idx = rng(41).permutation(k + 3)[:k].astype(np.int32)
    lam = rng(42).uniform(0.1, 10.0, k)
    mA = Matrix(A)
    f32 = case.endswith("f32_k9")
    x = MultiVec(mA, k + 3, F32, data=X) if f32 else MultiVec(mA, k + 3, data=X)
    if f32:
        X = X.astype(np.float32).astype(np.float64)
    rn, xn, bn = ops.residual_norms(mA, None, x, idx, lam)
    R = A @ X[:, idx] - X[:, idx] * lam[None, :]
    np.testing.assert_allclose(rn, np.linalg.norm(R, axis=0), rtol=1e-11)
    np.testing.assert_allclose(xn, np.linalg.norm(X[:, idx], axis=0), rtol=1e-12)
    np.testing.assert_allclose(bn, xn, rtol=0)


@pytest.mark.parametrize("name", ["vardiff2d", "chunglu"])
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("k", [2, 5, 17, 70, 130])
def test_residual_norms_row_gather(ops, monkeypatch, name, prec, k):
    # the row-major gather path (B = I): same FP64 row sums as the column
    # path, so the norms agree with it and with numpy to summation order
    A = MATS[name]
    n = A.shape[0]
    X = rng(k + 50).standard_normal((n, k + 3))
    idx = rng(k + 51).permutation(k + 3)[:k].astype(np.int32)
    lam = rng(k + 52).uniform(0.1, 10.0, k)
    res = []
    for layout in ("rows", "columns"):
        mA = _matrix_with_gather(A, layout, monkeypatch)
        x = MultiVec(mA, k + 3, prec, data=X)
        res.append(ops.residual_norms(mA, None, x, idx, lam))
    if prec == F32:
        X = X.astype(np.float32).astype(np.float64)
    R = A @ X[:, idx] - X[:, idx] * lam[None, :]
    np.testing.assert_allclose(res[0][0], np.linalg.norm(R, axis=0), rtol=1e-11)
    np.testing.assert_allclose(res[0][1], np.linalg.norm(X[:, idx], axis=0), rtol=1e-12)
    np.testing.assert_allclose(res[0][2], res[0][1], rtol=0)
    for a, b in zip(res[0], res[1]):
        np.testing.assert_allclose(a, b, rtol=1e-13)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("k", [3, 40])
@pytest.mark.parametrize("layout", ["rows", "columns"])
def test_residual_norms_row_gather_with_b(ops, monkeypatch, prec, k, layout):
    # generalized problem: B (27-point mass) on the row-gather path makes the
    # residual check gather both A's and B's rows from the staged X rows
    A, B = P.lap3d_mass(9)
    n = A.shape[0]
    X = rng(k + 60).standard_normal((n, k + 2))
    idx = rng(k + 61).permutation(k + 2)[:k].astype(np.int32)
    lam = rng(k + 62).uniform(1.0, 100.0, k)
    mA = _matrix_with_gather(A, layout, monkeypatch)
    mB = _matrix_with_gather(B, "rows", monkeypatch)
    x = MultiVec(mA, k + 2, prec, data=X)
    rn, xn, bn = ops.residual_norms(mA, mB, x, idx, lam)
    if prec == F32:
        X = X.astype(np.float32).astype(np.float64)
    BX = B @ X[:, idx]
    R = A @ X[:, idx] - BX * lam[None, :]
    np.testing.assert_allclose(rn, np.linalg.norm(R, axis=0), rtol=1e-11)
    np.testing.assert_allclose(xn, np.linalg.norm(X[:, idx], axis=0), rtol=1e-12)
    np.testing.assert_allclose(bn, np.linalg.norm(BX, axis=0), rtol=1e-12)
    # and the column path of the same pair agrees to summation order
    mB2 = _matrix_with_gather(B, "columns", monkeypatch)
    mA2 = _matrix_with_gather(A, "columns", monkeypatch)
    x2 = MultiVec(mA2, k + 2, prec, data=X)
    ref = ops.residual_norms(mA2, mB2, x2, idx, lam)
    for a, b in zip((rn, xn, bn), ref):
        np.testing.assert_allclose(a, b, rtol=1e-13)


def test_residual_norms_fp32_storage_fp64_math(ops):
    A, _ = P.lap3d_mass(9)
    n, k = A.shape[0], 4
    X = rng(26).standard_normal((n, k))
    X32 = X.astype(np.float32).astype(np.float64)
    lam = np.array([1.0, 2.0, 3.0, 4.0])
    mA = Matrix(A)
    x = MultiVec(mA, k, F32, data=X)
    rn, xn, _ = ops.residual_norms(mA, None, x, np.arange(k), lam)
    R = A @ X32 - X32 * lam[None, :]
    np.testing.assert_allclose(rn, np.linalg.norm(R, axis=0), rtol=1e-11)


def test_gather_copy_swap(ops):
    A = MATS["vardiff2d"]
    n = A.shape[0]
    X = rng(27).standard_normal((n, 10))
    Y = rng(28).standard_normal((n, 10))
    mA = Matrix(A)
    x = MultiVec(mA, 10, data=X)
    y = MultiVec(mA, 10, data=Y)
    ops.gather(x, [7, 2, 2, 9], y, ys=3)
    ref = Y.copy()
    ref[:, 3:7] = X[:, [7, 2, 2, 9]]
    np.testing.assert_array_equal(y.get(), ref)
    ops.swap(x, y, 1, 4, 3)
    rx, ry = X.copy(), ref.copy()
    rx[:, 1:4], ry[:, 4:7] = ref[:, 4:7], X[:, 1:4]
    np.testing.assert_array_equal(x.get(), rx)
    np.testing.assert_array_equal(y.get(), ry)
    # swap within one multivector (deflation moves a column to the end)
    ops.swap(x, x, 2, 9, 1)
    rx[:, [2, 9]] = rx[:, [9, 2]]
    np.testing.assert_array_equal(x.get(), rx)
    # precision-converting copy
    z = MultiVec(mA, 10, F32)
    ops.copy(x, z)
    np.testing.assert_array_equal(z.get(), rx.astype(np.float32).astype(np.float64))


def test_dirichlet_rows(ops):
    """SetDirichletBoundary zeroes rows whose B row is empty and whose A row
    has one entry (CSR_SetDirichletBoundary, gcge_app_csr.c:130-183)."""
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "A5M5_example.npz"))
    n = int(d["n"])
    A = sp.csr_matrix((d["A_data"], d["A_indices"], d["A_indptr"]), shape=(n, n))
    B = sp.csr_matrix((d["B_data"], d["B_indices"], d["B_indptr"]), shape=(n, n))
    X = rng(29).standard_normal((n, 4))
    mA, mB = Matrix(A), Matrix(B)
    x = MultiVec(mA, 4, data=X)
    ops.dirichlet(x, 4, mA, mB)
    bz = np.array([np.all(np.abs(B.data[B.indptr[i]:B.indptr[i + 1]]) <= 1e-20) for i in range(n)])
    a1 = np.array([np.sum(np.abs(A.data[A.indptr[i]:A.indptr[i + 1]]) > 1e-20) == 1 for i in range(n)])
    ref = X.copy()
    ref[bz & a1, :] = 0.0
    np.testing.assert_array_equal(x.get(), ref)


def test_range_errors_are_latched_not_launched(ops):
    A = MATS["vardiff2d"]
    mA = Matrix(A)
    x = MultiVec(mA, 3)
    y = MultiVec(mA, 3)
    with pytest.raises(RuntimeError, match="outside"):
        ops.matdot(mA, x, y, xs=1, k=3)
    # the latch was cleared by the raise; the backend keeps working
    ops.matdot(mA, x, y)


def test_sampled_outer_gradient_kernel():
    """grad_p = sum_i w_i x_i^T (dA/dtheta_p) x_i with dA/dtheta_p given per
    stored entry: compared with the explicit sparse product on the CPU."""
    A = MATS["vardiff2d"]
    n, k = A.shape[0], 5
    X = rng(30).standard_normal((n, k))
    w = rng(31).uniform(0.5, 2.0, k)
    nparam = 37
    g = rng(32)
    cnt = g.integers(0, 6, nparam)
    jrp = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    jnz = g.integers(0, A.nnz, jrp[-1]).astype(np.int32)
    jval = g.standard_normal(jrp[-1])
    mA = Matrix(A)
    grad = eig_gradient(mA, X, w, nparam, jrp, jnz, jval)
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    ge = np.zeros(A.nnz)
    for i in range(k):
        ge += w[i] * X[rows, i] * X[A.indices, i]
    ref = np.array([np.sum(jval[jrp[p]:jrp[p + 1]] * ge[jnz[jrp[p]:jrp[p + 1]]]) for p in range(nparam)])
    np.testing.assert_allclose(grad, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
