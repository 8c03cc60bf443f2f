"""Writes tests/golden/large/chunglu200k_ref.npz by running the real
reference (oracle/_ref/ref_solve, built from the reference's own sources with
`make -C oracle ref`) on a power-law graph too large to store as a matrix.

The case: graph Laplacian of a Chung-Lu random graph (problems.chung_lu,
n = 200,000, exponent 2.5, mean degree 16, seed 0) plus 1e-2 I, B = I, the
lowest 64 eigenpairs from X0 = problems.random_x0(n, 64, 1), ev_tol 1e-8 with
the reference's relative criterion "R", at most 250 outer iterations.  The
graph has 132 isolated vertices, so all 64 wanted eigenvalues are 1e-2: the
subspace eigensolver works inside a degenerate cluster from the first step.
The fixture keeps the generator parameters, a fingerprint of the matrix and
the reference's outputs (eigenvalues, residuals, iteration count, per-
iteration history); the test regenerates matrix and X0.

Run (in a container that has /root/reference; about 30 minutes on one core):
    MKL_NUM_THREADS=1 python tests/golden/large/make_chunglu200k.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "gcge-1.0_amd"))

import oracle_api as O  # noqa: E402
from gcge_hip import problems as P  # noqa: E402

N, GRAPH_SEED, NEV, X0_SEED = 200_000, 0, 64, 1
FLAGS = ["-gcge_ev_max_it", "250", "-gcge_ev_tol", "1e-8", "-gcge_conv_type", "R", "-gcge_print_para", "0"]


def fingerprint(A):
    return np.array([A.shape[0], A.nnz, int(A.indptr.astype(np.int64).sum()),
                     int(A.indices.astype(np.int64).sum())], dtype=np.int64), float(A.data.sum())


def main():
    os.environ["MKL_NUM_THREADS"] = "1"
    if not O.ref_available():
        O.build_ref()
    A, _ = P.chung_lu(N, seed=GRAPH_SEED)
    X0 = P.random_x0(N, NEV, X0_SEED)
    r = O.ref_solve(A, None, NEV, X0=X0, flags=FLAGS, timeout=3600)
    its, _ = O.parse_trace(r["stdout"])
    ints, dsum = fingerprint(A)
    np.savez_compressed(
        os.path.join(HERE, "chunglu200k_ref.npz"),
        n=N, graph_seed=GRAPH_SEED, nev=NEV, x0_seed=X0_SEED, flags=np.array(FLAGS),
        fp_ints=ints, fp_data_sum=dsum,
        eval=r["eval"], res=r["res"], num_iter=r["info"]["num_iter"], num_unlock=r["info"]["num_unlock"],
        hist_eval=np.array([it["eval"] for it in its]), hist_res=np.array([it["res"] for it in its]),
        hist_unlock=np.array([it["num_unlock"] for it in its], dtype=np.int32))
    print(f"chunglu200k: iters={r['info']['num_iter']} unconv={r['info']['num_unlock']} "
          f"lambda range {r['eval'].min():.15e} .. {r['eval'].max():.15e}")


if __name__ == "__main__":
    main()
