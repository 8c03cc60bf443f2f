"""Writes the golden fixtures in tests/golden/*.npz by running the real
reference (oracle/_ref/ref_solve, built from the reference's own sources with
`make -C oracle ref`).

Each fixture holds the inputs (matrices, X0 or the rand() seed, parameters)
and the reference's outputs: eigenvalues, residuals, eigenvectors, iteration
counts and the per-iteration eigenvalue/residual history it prints.  The CSR
matrices A_3/M_3, A_5/M_5 and testA/testB are the data files of the
reference's own tests (test/data/), stored here as arrays.

Run (in a container that has /root/reference; names select fixtures):
    MKL_NUM_THREADS=1 python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "gcge-1.0_amd"))

import oracle_api as O  # noqa: E402
from gcge_hip import problems as P  # noqa: E402

REF_DATA = "/root/reference/test/data"


def read_csr_text(path):
    """CSR text format of CSR_ReadMatFile (app/csr/gcge_app_csr.c:185-236)."""
    with open(path) as f:
        toks = f.read().split()
    n, _, nnz = int(toks[0]), int(toks[1]), int(toks[2])
    ia = np.array(toks[3:4 + n], dtype=np.int32)
    ja = np.array(toks[4 + n:4 + n + nnz], dtype=np.int32)
    aa = np.array(toks[4 + n + nnz:4 + n + 2 * nnz], dtype=np.float64)
    return sp.csr_matrix((aa, ja, ia), shape=(n, n))


def cases():
    A3, M3 = read_csr_text(f"{REF_DATA}/A_3.txt"), read_csr_text(f"{REF_DATA}/M_3.txt")
    A5, M5 = read_csr_text(f"{REF_DATA}/A_5.txt"), read_csr_text(f"{REF_DATA}/M_5.txt")
    tA, tB = read_csr_text(f"{REF_DATA}/testA"), read_csr_text(f"{REF_DATA}/testB")
    L16, _ = P.lap2d(16)
    L32, _ = P.lap2d(32)
    A8, B8 = P.lap3d_mass(8)
    V16, _ = P.vardiff2d(16, seed=3)
    # (name, A, B, nev, X0, seed, dirichlet, flags, orc_kwargs)
    out = [
        ("lap2d16_rand", L16, None, 6, None, 7, 0, [], {}),
        ("A3M3_rand", A3, M3, 5, None, 7, 0, [], {}),
        # example/csr_solver.c:41-61: A_5/M_5, nev=10, Dirichlet rows
        ("A5M5_example", A5, M5, 10, None, 7, 1, [], dict(dirichlet_boundary=1)),
        # test/csr/test_csr_solver.c:38-56: nev=20, cg_max_it 20, ev_max_it 150, cg_type 2
        ("A5M5_testsolver", A5, M5, 20, None, 11, 1,
         ["-gcge_cg_max_it", "20", "-gcge_ev_max_it", "150", "-cg_type", "2"],
         dict(dirichlet_boundary=1, cg_max_it=20, ev_max_it=150, cg_type=2)),
        ("lap2d32_x0", L32, None, 20, P.random_x0(L32.shape[0], 20, 5), 1, 0,
         ["-gcge_ev_tol", "1e-10", "-gcge_ev_max_it", "80"], dict(ev_tol=1e-10, ev_max_it=80)),
        ("lap3dmass8_x0", A8, B8, 10, P.random_x0(A8.shape[0], 10, 6), 1, 0,
         ["-gcge_ev_max_it", "60"], dict(ev_max_it=60)),
        ("vardiff2d16_x0", V16, None, 8, P.random_x0(V16.shape[0], 8, 8), 1, 0,
         ["-gcge_ev_max_it", "60"], dict(ev_max_it=60)),
        # scalar CG (if_use_bcg = 0) and plain Gram-Schmidt for W
        ("lap2d16_scalarcg", L16, None, 6, P.random_x0(L16.shape[0], 6, 9), 1, 0,
         ["-gcge_if_use_bcg", "0", "-cg_type", "2"], dict(if_use_bcg=0, cg_type=2)),
        ("lap2d16_scalarcg1", L16, None, 6, P.random_x0(L16.shape[0], 6, 9), 1, 0,
         ["-gcge_if_use_bcg", "0"], dict(if_use_bcg=0, cg_type=1)),
        # cg_rate 1e-3 (set directly: the reference parses -gcge_cg_rate with atoi)
        ("lap2d16_cgrate", L16, None, 6, P.random_x0(L16.shape[0], 6, 14), 1, 0,
         ["-cg_rate", "1e-3"], dict(cg_rate=1e-3)),
        ("lap2d16_gsorth", L16, None, 6, P.random_x0(L16.shape[0], 6, 10), 1, 0,
         ["-gcge_w_orth_type", "gs"], dict(w_orth_type="gs")),
        # unconverged exit after one outer step
        ("lap2d16_maxit1", L16, None, 6, P.random_x0(L16.shape[0], 6, 12), 1, 0,
         ["-gcge_ev_max_it", "1"], dict(ev_max_it=1)),
        # 3x3 SPD matrix of test/data/testA with B = I3 from testB
        ("testA_nev1", tA, tB, 1, None, 3, 0, [], {}),
    ]
    # the other W orthonormalizations of the reference (gcge_eigsol.c:448-463),
    # standard and generalized
    for wt in ("bgs", "cbgs", "scbgs"):
        out.append((f"lap2d16_{wt}", L16, None, 6, P.random_x0(L16.shape[0], 6, 15), 1, 0,
                    ["-gcge_w_orth_type", wt], dict(w_orth_type=wt)))
        out.append((f"lap3dmass8_{wt}", A8, B8, 10, P.random_x0(A8.shape[0], 10, 16), 1, 0,
                    ["-gcge_w_orth_type", wt, "-gcge_ev_max_it", "60"], dict(w_orth_type=wt, ev_max_it=60)))
    # the in-subspace orthonormalizations of ComputeP's coefficients (gcge_xpw.c:363-385)
    for pt in ("gs", "bgs", "scbgs"):
        out.append((f"lap2d16_p{pt}", L16, None, 6, P.random_x0(L16.shape[0], 6, 20), 1, 0,
                    ["-gcge_p_orth_type", pt], dict(p_orth_type=pt)))
        out.append((f"lap3dmass8_p{pt}", A8, B8, 10, P.random_x0(A8.shape[0], 10, 21), 1, 0,
                    ["-gcge_p_orth_type", pt, "-gcge_ev_max_it", "60"], dict(p_orth_type=pt, ev_max_it=60)))
    # scbgs with W in blocks of 2 columns (w_orth_block_size) and two MGS sweeps per block
    out.append(("lap2d32_scbgs_blk", L32, None, 12, P.random_x0(L32.shape[0], 12, 17), 1, 0,
                ["-gcge_w_orth_type", "scbgs", "-gcge_w_orth_block_size", "2",
                 "-gcge_scbgs_wself_max_reorth_time", "2", "-gcge_ev_max_it", "80"],
                dict(w_orth_type="scbgs", ev_max_it=80, w_orth_block_size=2, scbgs_wself_max_reorth_time=2)))
    # soft locking: Rayleigh-Ritz pairs from the first unconverged one on
    # (opt_rr_eig_partly, gcge_rayleighritz.c:380-394)
    out.append(("lap2d32_rrpartly", L32, None, 20, P.random_x0(L32.shape[0], 20, 18), 1, 0,
                ["-gcge_opt_rr_eig_partly", "1", "-gcge_ev_tol", "1e-10", "-gcge_ev_max_it", "80"],
                dict(opt_rr_eig_partly=1, ev_tol=1e-10, ev_max_it=80)))
    out.append(("lap3dmass8_rrpartly", A8, B8, 10, P.random_x0(A8.shape[0], 10, 19), 1, 0,
                ["-gcge_opt_rr_eig_partly", "1", "-gcge_ev_max_it", "60"],
                dict(opt_rr_eig_partly=1, ev_max_it=60)))
    # rank-deficient X0: duplicated columns exercise the deflation swaps
    X = P.random_x0(L16.shape[0], 6, 13)
    X[:, 3] = X[:, 1]
    out.append(("lap2d16_dupx0", L16, None, 6, X, 1, 0, [], {}))
    # a zero column: deflated for certain at the initial orthonormalization
    X = P.random_x0(L16.shape[0], 6, 13)
    X[:, 3] = 0.0
    out.append(("lap2d16_zerox0", L16, None, 6, X, 1, 0, [], {}))
    return out


def main(names=()):
    os.environ["MKL_NUM_THREADS"] = "1"
    if not O.ref_available():
        O.build_ref()
    for name, A, B, nev, X0, seed, dirich, flags, okw in cases():
        if names and name not in names:
            continue
        r = O.ref_solve(A, B, nev, X0=X0, flags=flags, seed=seed, dirichlet=dirich)
        its, notices = O.parse_trace(r["stdout"])
        hist_ev = np.array([it["eval"] for it in its]) if its else np.zeros((0, nev))
        hist_res = np.array([it["res"] for it in its]) if its else np.zeros((0, nev))
        hist_unlock = np.array([it["num_unlock"] for it in its], dtype=np.int32)
        d = dict(
            n=A.shape[0], nev=nev, seed=seed, dirichlet=dirich,
            A_indptr=A.indptr.astype(np.int32), A_indices=A.indices.astype(np.int32),
            A_data=A.data.astype(np.float64),
            has_B=B is not None, has_X0=X0 is not None,
            eval=r["eval"], res=r["res"], evec=r["evec"],
            num_iter=r["info"]["num_iter"], num_unlock=r["info"]["num_unlock"],
            hist_eval=hist_ev, hist_res=hist_res, hist_unlock=hist_unlock,
            notices=notices,
            orc_kwargs=np.array(repr(okw)),
        )
        if B is not None:
            d.update(B_indptr=B.indptr.astype(np.int32), B_indices=B.indices.astype(np.int32),
                     B_data=B.data.astype(np.float64))
        if X0 is not None:
            d["X0"] = np.asfortranarray(X0)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **d)
        print(f"{name}: n={A.shape[0]} nev={nev} iters={r['info']['num_iter']} "
              f"unconv={r['info']['num_unlock']} lambda0={r['eval'][0]:.12e}")


if __name__ == "__main__":
    main(sys.argv[1:])
