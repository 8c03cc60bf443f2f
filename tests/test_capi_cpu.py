"""CPU checks of the C ABI (no GPU needed): the library loads, exports every
function declared in include/*.h, its struct layouts match the bindings, and
the host-side parts (parameters, option parsing, ops-table setup, the
multiplicity rule) behave as the reference's."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from gcge_hip import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = capi.LIB_PATH


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        # drop struct bodies (function-pointer members are not exports)
        txt = re.sub(r"typedef struct[^{;]*\{.*?\}\s*\w+\s*;", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b((?:GCGE|HIP)_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    decl = declared_functions()
    assert len(decl) > 55
    missing = sorted(decl - exported)
    assert not missing, missing
    # and the ctypes binding knows all of them
    assert not sorted(decl - set(capi.SIGNATURES)), sorted(decl - set(capi.SIGNATURES))


def test_library_loads_and_binds():
    L = capi.load()
    for name in capi.exported_symbols():
        assert getattr(L, name) is not None


def test_abi_layout_matches_bindings():
    lib_side, py_side = capi.abi_check()
    assert lib_side == py_side


def test_ops_table_has_reference_slot_order():
    # the first 36 members are the reference's GCGE_OPS slots in order
    # (src/gcge_ops.h:29-215); the block-kernel slots follow
    names = [f for f, _ in capi.GCGE_OPS._fields_]
    assert names[:8] == ["VecSetRandomValue", "MatDotVec", "VecAxpby", "VecInnerProd", "VecLocalInnerProd",
                         "VecCreateByVec", "VecCreateByMat", "VecDestroy"]
    assert names[35] == "ArrayScale"
    assert capi.GCGE_OPS.ArrayScale.offset == 35 * 8


def test_para_defaults_match_reference():
    L = capi.load()
    p = C.POINTER(capi.GCGE_PARA)()
    L.GCGE_PARA_Create(C.byref(p))
    P = p.contents
    # src/gcge_para.c:25-99
    assert (P.nev, P.ev_max_it, P.block_size, P.given_init_evec) == (1, 30, 0, 0)
    assert P.ev_tol == 1e-8 and P.conv_type == b"R" and P.orth_type == b"B"
    assert P.w_orth_type == b"multi" and P.p_orth_type == b"multi"
    o = P.orth_para.contents
    assert (o.orth_zero_tol, o.reorth_tol, o.scbgs_reorth_tol, o.criterion_tol) == (1e-16, 0.75, 1e-16, 1e-6)
    assert (o.max_reorth_time, o.max_direct_orth_length) == (3, 1)
    assert (P.cg_max_it, P.cg_rate, P.cg_type, P.if_use_bcg) == (20, 1e-2, 1, 1)
    assert P.multi_tol == 0.2 and P.multi_tol_for_lock == 1e-6
    args = [b"prog", b"-gcge_nev", b"12", b"-gcge_ev_tol", b"1e-10", b"-gcge_cg_rate", b"1e-3",
            b"-gcge_w_orth_type", b"gs", b"-unknown", b"-gcge_block_size", b"4"]
    argv = (C.c_char_p * len(args))(*args)
    assert L.GCGE_PARA_SetFromCommandLine(p, len(args), argv) == 0
    assert P.nev == 12 and P.ev_tol == 1e-10 and P.block_size == 4 and P.w_orth_type == b"gs"
    assert P.cg_rate == 1e-3
    L.GCGE_PARA_Setup(p)
    assert P.num_unlock == 12
    assert o.orth_zero_tol == 1e-16 + np.finfo(float).eps
    assert [P.res[i] for i in range(12)] == [1.0] * 12
    L.GCGE_PARA_Free(C.byref(p))


def test_ops_setup_rejects_missing_block_kernels():
    L = capi.load()
    ops = C.POINTER(capi.GCGE_OPS)()
    L.GCGE_OPS_Create(C.byref(ops))
    assert L.GCGE_OPS_Setup(ops) == 1          # empty table: refused
    L.GCGE_HIP_SetOps(ops)
    assert L.GCGE_OPS_Setup(ops) == 0          # HIP backend fills every slot
    for name, _ in capi.GCGE_OPS._fields_:
        if name not in ("LinearSolver", "linear_solver_workspace", "Orthonormalization", "DenseMatCreate",
                        "DenseMatDestroy"):
            assert getattr(ops.contents, name), name
    L.GCGE_OPS_Free(C.byref(ops))


@pytest.mark.parametrize("ev,exp", [
    ([1.0, 1.1, 1.15, 1.3, 2.0, 2.1], 4),    # nev=3: 1.15, 1.3 within 0.2 of eval[3]=1.3? ratio rule
    ([1.0, 2.0, 3.0, 3.5, 3.55, 10.0], 5),
])
def test_check_eval_multiplicity_rule(ev, exp):
    """dim_x = nev + #consecutive i >= nev with |eval[i]/eval[nev] - 1| < 0.2
    (gcge_eigsol.c:696-718), checked against a direct evaluation."""
    L = capi.load()
    nev, end = 3, 6
    ev = np.array(ev)
    d = C.c_int(0)
    L.GCGE_CheckEvalMultiplicity(nev, nev, end, C.byref(d), ev.ctypes.data_as(capi.DP))
    t = nev
    for i in range(nev, end):
        if abs(abs(ev[t] / ev[nev]) - 1) < 0.2:
            t += 1
        else:
            break
    assert d.value == t


def test_no_gpu_calls_fail_cleanly_without_device():
    """Without a GPU (this container) device entry points latch an error and
    return NULL instead of crashing."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    L = capi.load()
    rp = np.array([0, 1], dtype=np.int32)
    ci = np.array([0], dtype=np.int32)
    v = np.array([2.0])
    p = L.HIP_MatCreateFromCSR(1, 1, 1, rp.ctypes.data_as(capi.IP), ci.ctypes.data_as(capi.IP),
                               v.ctypes.data_as(capi.DP))
    assert not p
    code, msg = capi.last_error()
    assert code != 0 and msg
    L.GCGE_HIP_ClearError()
    assert capi.last_error()[0] == 0


def _dense_eig_slot():
    L = capi.load()
    ops = capi.GCGE_OPS()
    L.GCGE_SetDenseOps.argtypes = [C.POINTER(capi.GCGE_OPS)]
    L.GCGE_SetDenseOps.restype = None
    L.GCGE_SetDenseOps(C.byref(ops))
    I, D = C.POINTER(C.c_int), C.POINTER(C.c_double)
    proto = C.CFUNCTYPE(None, C.c_char_p, C.c_char_p, C.c_char_p, I, D, I, D, D, I, I, D, I, D, D, I, I, D, I, I,
                        I, I, I)
    return proto(ops.DenseMatEigenSolver)


@pytest.mark.parametrize("cluster", [0, 40, 150])
def test_dense_eigensolver_range_on_degenerate_clusters(cluster):
    # the subspace eigenproblem of the outer iteration (src/gcge_ops.c:302-336,
    # dsyevx range "I"): pairs il..iu of a symmetric matrix whose lowest
    # `cluster` eigenvalues are equal (a graph Laplacian plus a shift with
    # many isolated vertices gives such Ritz matrices); the pairs must be
    # the lowest eigenvalues and an orthonormal set of eigenvectors
    eig = _dense_eig_slot()
    n, il, iu = 208, 1, 108
    rng = np.random.default_rng(cluster + 7)
    lam = np.sort(np.concatenate([np.full(cluster, 1e-2), 1e-2 + rng.uniform(1e-3, 5.0, n - cluster)]))
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    a = np.asfortranarray(A.copy())
    m = iu - il + 1
    ev = np.zeros(n)  # dsyevx may use all n entries
    V = np.zeros((n, m), order="F")
    ci = lambda v: C.byref(C.c_int(v))
    nev, info = C.c_int(0), C.c_int(-99)
    lwork = 8 * n
    work, iwork, ifail = np.zeros(lwork), np.zeros(5 * n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    isuppz = np.zeros(2 * n, dtype=np.int32)
    dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda x: x.ctypes.data_as(C.POINTER(C.c_int))
    eig(b"V", b"I", b"U", ci(n), dp(a), ci(n), C.byref(C.c_double(0)), C.byref(C.c_double(0)), ci(il), ci(iu),
        C.byref(C.c_double(0)), C.byref(nev), dp(ev), dp(V), ci(n), ip(isuppz), dp(work), ci(lwork), ip(iwork),
        ci(5 * n), ip(ifail), C.byref(info))
    assert info.value == 0 and nev.value == m
    ev = ev[:m]
    np.testing.assert_allclose(ev, lam[il - 1:iu], rtol=1e-12, atol=1e-13)
    assert np.abs(V.T @ V - np.eye(m)).max() < 1e-12
    assert np.abs(A @ V - V * ev).max() < 1e-12 * np.abs(lam).max() * 10


def _conv_setup(nev, res_values):
    """GCGE_PARA (conv_type R, ev_tol 1e-8) and a workspace for
    GCGE_CheckConvergence, with an ops table whose MultiVecResidualNorms
    returns ||r|| = res_values[idx], ||x|| = 1 (no device involved)."""
    L = capi.load()
    p = C.POINTER(capi.GCGE_PARA)()
    L.GCGE_PARA_Create(C.byref(p))
    L.GCGE_PARA_SetNumEigen(p, nev)
    L.GCGE_PARA_Setup(p)
    ws = capi.GCGE_WORKSPACE()
    unlock = (C.c_int * nev)(*range(nev))
    ws.unlock = C.cast(unlock, capi.IP)
    ws.max_dim_x = nev + 3
    ops = capi.GCGE_OPS()
    calls = []

    def resnorms(A, B, X, idx, k, lam, rn, xn, bn, o):
        for t in range(k):
            calls.append(idx[t])
            rn[t] = res_values[idx[t]]
            xn[t] = 1.0
            if bn:
                bn[t] = 1.0

    cb = capi.F_MultiVecResidualNorms(resnorms)
    ops.MultiVecResidualNorms = C.cast(cb, C.c_void_p)
    keep = (unlock, cb)
    return L, p, ws, ops, calls, keep


def test_nan_residual_is_not_converged():
    """The reference tests residual > ev_tol (src/gcge_eigsol.c:630), so a
    NaN residual passes as converged; here it stays unconverged, becomes the
    largest residual and latches error 190."""
    nev = 6
    res = [1e-12, float("nan"), 1e-12, 1e-12, float("inf"), 1e-12]
    L, p, ws, ops, calls, keep = _conv_setup(nev, res)
    ev = np.arange(1.0, nev + 1.0) * 0.1
    L.GCGE_HIP_ClearError()
    L.GCGE_CheckConvergence(None, None, ev.ctypes.data_as(capi.DP), None, C.byref(ops), p, C.byref(ws))
    P = p.contents
    assert P.num_unlock == 2
    assert [ws.unlock[i] for i in range(2)] == [1, 4]
    assert np.isnan(P.max_res)
    code, msg = capi.last_error()
    assert code == 190 and "eigenpair 1" in msg
    L.GCGE_HIP_ClearError()
    L.GCGE_PARA_Free(C.byref(p))


def test_finite_residuals_unchanged_rule():
    """Finite residuals follow the reference's rule, including the early exit
    eight pairs past the last converged one (src/gcge_eigsol.c:637-645)."""
    nev = 14
    res = [1e-12] * 2 + [1.0] * 12
    L, p, ws, ops, calls, keep = _conv_setup(nev, res)
    ev = np.arange(1.0, nev + 1.0) * 0.1
    L.GCGE_HIP_ClearError()
    L.GCGE_CheckConvergence(None, None, ev.ctypes.data_as(capi.DP), None, C.byref(ops), p, C.byref(ws))
    P = p.contents
    assert P.num_unlock == 12
    assert capi.last_error()[0] == 0
    # pairs 0-1 converge (last converged position 1); the check stops at the
    # first unconverged position past 1 + 8, so pairs 11.. are never visited
    assert sorted(set(calls)) == list(range(11))
    assert P.num_iter == 1
    L.GCGE_PARA_Free(C.byref(p))


@pytest.mark.parametrize("nev,bs,max_dim_x", [
    (1, 1, 4),      # nev * 1.25 rounds to 1; the nev + 3 floor applies
    (10, 5, 13),    # 12 = int(12.5) < nev + 8, below the nev + 3 floor
    (20, 20, 25),   # int(25.0) < nev + 8 = 28
    (40, 40, 48),   # nev + 8 = 48 < 50
    (100, 100, 108),
    (100, 36, 108),
])
def test_workspace_sizes_follow_reference_rule(nev, bs, max_dim_x):
    """GCGE_WORKSPACE_Setup requests the multivectors gcge_workspace.c:61-70
    sizes: max_dim_x = max(min(int(1.25 nev), nev + 8), nev + 3), V of
    max_dim_x + 2 bs columns, V_tmp of max_dim_x - nev + bs, CG_p of bs.  The
    expected max_dim_x values are worked out by hand for each branch of that
    rule; the ops table records the requests (no device involved)."""
    L = capi.load()
    p = C.POINTER(capi.GCGE_PARA)()
    L.GCGE_PARA_Create(C.byref(p))
    L.GCGE_PARA_SetNumEigen(p, nev)
    p.contents.block_size = bs
    L.GCGE_PARA_Setup(p)
    assert p.contents.nev == nev and p.contents.block_size == bs
    ops = capi.GCGE_OPS()
    requests, destroyed = [], []
    handles = [C.c_int(0) for _ in range(3)]

    F_create = C.CFUNCTYPE(None, C.POINTER(C.c_void_p), C.c_int, C.c_void_p, C.c_void_p)
    F_destroy = C.CFUNCTYPE(None, C.POINTER(C.c_void_p), C.c_int, C.c_void_p)

    def create(out, k, A, o):
        out[0] = C.cast(C.byref(handles[len(requests)]), C.c_void_p).value
        requests.append(k)

    def destroy(h, k, o):
        destroyed.append(k)
        h[0] = None

    cbc, cbd = F_create(create), F_destroy(destroy)
    ops.MultiVecCreateByMat = C.cast(cbc, C.c_void_p)
    ops.MultiVecDestroy = C.cast(cbd, C.c_void_p)
    ws = C.POINTER(capi.GCGE_WORKSPACE)()
    L.GCGE_WORKSPACE_Create(C.byref(ws))
    L.GCGE_WORKSPACE_Setup(ws, p, C.byref(ops), None)
    W = ws.contents
    assert requests == [max_dim_x + 2 * bs, max_dim_x - nev + bs, bs]
    assert (W.V_size, W.V_tmp_size, W.CG_p_size) == tuple(requests)
    assert W.max_dim_x == max_dim_x
    assert (W.dim_x, W.dim_xp, W.dim_xpw, W.conv_bs) == (nev, 0, nev, 0)
    assert W.unconv_bs == min(nev, bs)
    assert [W.unlock[i] for i in range(max_dim_x)] == list(range(max_dim_x))
    L.GCGE_WORKSPACE_Free(C.byref(ws), p, C.byref(ops))
    assert destroyed == requests
    L.GCGE_PARA_Free(C.byref(p))


def test_w_orth_types_parse():
    """-gcge_w_orth_type accepts the reference's names (multi, gs, bgs, cbgs,
    scbgs) and this build's cholqr."""
    L = capi.load()
    for wt in (b"multi", b"gs", b"bgs", b"cbgs", b"scbgs", b"cholqr"):
        p = C.POINTER(capi.GCGE_PARA)()
        L.GCGE_PARA_Create(C.byref(p))
        args = [b"prog", b"-gcge_w_orth_type", wt]
        argv = (C.c_char_p * len(args))(*args)
        assert L.GCGE_PARA_SetFromCommandLine(p, len(args), argv) == 0
        assert p.contents.w_orth_type == wt
        L.GCGE_PARA_Free(C.byref(p))


def _read_text_csr(path):
    """The reference's reader (CSR_ReadMatFile, app/csr/gcge_app_csr.c:185-236):
    nrow, ncol, nnz, then rowptr, col, val, whitespace separated."""
    toks = open(path).read().split()
    n, m, nnz = int(toks[0]), int(toks[1]), int(toks[2])
    ia = np.array(toks[3:4 + n], dtype=np.int64)
    ja = np.array(toks[4 + n:4 + n + nnz], dtype=np.int64)
    aa = np.array([float(t) for t in toks[4 + n + nnz:4 + n + 2 * nnz]])
    return n, m, nnz, ia, ja, aa


@pytest.mark.parametrize("binary", [0, 1])
def test_write_mat_file_formats(tmp_path, binary):
    """HIP_WriteMatFile (host only): the text form is the reference's reader
    format with values that read back exactly; the binary form is the
    "GCGECSR1" header and the three arrays."""
    from gcge_hip import problems as P
    from gcge_hip import write_matrix_file
    A, B = P.lap3d_mass(5)
    path = tmp_path / ("B.bin" if binary else "B.txt")
    write_matrix_file(path, B, binary=bool(binary))
    if not binary:
        n, m, nnz, ia, ja, aa = _read_text_csr(path)
    else:
        raw = path.read_bytes()
        assert raw[:8] == b"GCGECSR1"
        n, m, nnz, zero = np.frombuffer(raw[8:24], dtype="<i4")
        assert zero == 0
        off = 24
        ia = np.frombuffer(raw[off:off + 4 * (n + 1)], dtype="<i4")
        off += 4 * (n + 1)
        ja = np.frombuffer(raw[off:off + 4 * nnz], dtype="<i4")
        off += 4 * nnz
        aa = np.frombuffer(raw[off:off + 8 * nnz], dtype="<f8")
        assert off + 8 * nnz == len(raw)
    assert (n, m, nnz) == (B.shape[0], B.shape[1], B.nnz)
    np.testing.assert_array_equal(ia, B.indptr)
    np.testing.assert_array_equal(ja, B.indices)
    np.testing.assert_array_equal(aa, B.data)   # bit for bit


def _block_extras(M, perm, R=256):
    """Distinct rows outside each 256-row block that the block's entries read,
    for the matrix in the row order perm (perm[new] = row)."""
    n = M.shape[0]
    new = np.empty(n, dtype=np.int64)
    new[perm] = np.arange(n)
    out = []
    for b in range(0, n, R):
        rows = perm[b:b + R]
        cols = np.concatenate([new[M.indices[M.indptr[r]:M.indptr[r + 1]]] for r in rows])
        u = np.unique(cols)
        out.append(np.sum((u < b) | (u >= b + R)))
    return np.array(out)


@pytest.mark.parametrize("dims,tile", [((13, 11, 9), (8, 8, 8)), ((40, 24, 1), (16, 16, 1)),
                                       ((17, 16, 16), (8, 8, 8))])
def test_locality_order_tiles_a_box_grid(dims, tile):
    """A lexicographically numbered box grid (7-point in 3-D, 5-point in
    2-D; extents not multiples of the tile) comes back tile by tile: the
    expected order is built here from grid coordinates, tiles clipped at
    the faces, tiles and rows within a tile x fastest."""
    import scipy.sparse as sp
    from gcge_hip import locality_order
    nx, ny, nz = dims
    tx, ty, tz = tile

    def lap1(m):
        return sp.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])

    I = lambda m: sp.identity(m)
    if nz > 1:
        A = sp.kron(sp.kron(I(nz), I(ny)), lap1(nx)) + sp.kron(sp.kron(I(nz), lap1(ny)), I(nx)) + \
            sp.kron(sp.kron(lap1(nz), I(ny)), I(nx))
    else:
        A = sp.kron(I(ny), lap1(nx)) + sp.kron(lap1(ny), I(nx))
    A = A.tocsr()
    A.sort_indices()
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    row = ((z * ny + y) * nx + x).ravel()
    key = np.lexsort(((x % tx).ravel(), (y % ty).ravel(), (z % tz).ravel(),
                      (x // tx).ravel(), (y // ty).ravel(), (z // tz).ravel()))
    np.testing.assert_array_equal(locality_order(A), row[key])


def test_locality_order_is_a_permutation_with_compact_blocks():
    """HIP_MatLocalityOrder (host only): a permutation of the rows; on a 3-D
    7-point grid its 256-row blocks read about their own size in other rows
    (their surface), against three times that in the natural order (whole
    neighbouring planes); a graph with isolated vertices and several
    components is still fully ordered."""
    from gcge_hip import locality_order
    from gcge_hip import problems as P
    A, _ = P.vardiff3d(24, seed=0)
    perm = locality_order(A)
    n = A.shape[0]
    assert np.array_equal(np.sort(perm), np.arange(n))
    nat = _block_extras(A, np.arange(n)).mean()
    loc = _block_extras(A, perm).mean()
    assert loc < 0.6 * nat, (loc, nat)
    G, _ = P.chung_lu(5000, mean_deg=2.0, seed=3)
    pg = locality_order(G, leaf=64)
    assert np.array_equal(np.sort(pg), np.arange(G.shape[0]))
    # a grid graph whose numbering is scrambled is no box in natural order:
    # the recursive halving still finds compact blocks
    rng = np.random.default_rng(0)
    q = rng.permutation(n)
    S = A[q][:, q].tocsr()
    ps = locality_order(S)
    assert np.array_equal(np.sort(ps), np.arange(n))
    assert _block_extras(S, ps).mean() < 0.6 * nat
    with pytest.raises(ValueError):
        bad = A.copy()
        bad.indices[0] = n + 5
        locality_order(bad)
