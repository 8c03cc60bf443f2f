"""Writes tests/golden/large/{lap2d128,vardiff2d512,lap3dmass96}_ref.npz by running the
real reference (oracle/_ref/ref_solve, built from the reference's own sources
with `make -C oracle ref`) at the full sizes of three benchmark workloads:

* lap2d128:   2-D 5-point Dirichlet Laplacian on the 128 x 128 interior grid
              (problems.lap2d(128), n = 16,384), lowest 20 eigenpairs;
* vardiff2d512: 2-D variable-coefficient diffusion on 512 x 512 cells
              (problems.vardiff2d(512), x-faces x1.3, theta = exp(N(0, 0.25)),
              seed 0, n = 262,144), lowest 16 eigenpairs.

Both start from X0 = problems.random_x0(n, nev, seed) and use the
reference's relative criterion "R" (||r|| / ||x|| / max(1, |lambda|)) with
ev_tol = 1e-8 * 0.95 * lambda_1: with B = I and every wanted eigenvalue below
1, R <= ev_tol implies ||A x - lambda x|| / (|lambda| ||x||) <= 1e-8 for every
pair.  For the generalized case (lap3dmass96: 3-D Laplacian with the P1
mass matrix on 96^3, n = 884,736, lowest 50 pairs) every wanted eigenvalue
is above 1 and ev_tol = 1e-8 * 0.95 * min ||Bx|| / ||x|| over the wanted
modes (closed form), which again implies NS <= 1e-8.  lambda_1 comes from
scipy's shift-invert Lanczos before the run.  The fixture keeps the generator parameters, a fingerprint of the matrix and the
reference's outputs (eigenvalues, residuals, iteration count, per-iteration
history); the test regenerates matrix and X0.

Run (in a container that has /root/reference; minutes on one core for the
2-D cases, about an hour for lap3dmass96):
    MKL_NUM_THREADS=1 python tests/golden/large/make_fullsize.py [name ...]
"""
from __future__ import annotations

import os
import sys

import numpy as np
import scipy.sparse.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "gcge-1.0_amd"))

import oracle_api as O  # noqa: E402
from gcge_hip import problems as P  # noqa: E402

CASES = {
    "lap2d128": dict(gen="lap2d", m=128, gen_seed=0, nev=20, x0_seed=5),
    "vardiff2d512": dict(gen="vardiff2d", m=512, gen_seed=0, nev=16, x0_seed=6),
    # the generalized workload: 3-D Laplacian with the P1 mass matrix on 96^3
    # (n = 884,736), lowest 50 pairs
    "lap3dmass96": dict(gen="lap3d_mass", m=96, gen_seed=0, nev=50, x0_seed=7),
}


def mass_floor(m, nev):
    """Smallest ||Bx|| / ||x|| over the lowest nev modes of lap3d_mass(m):
    the modes are products of sines, eigenvectors of B as well, with
    B-eigenvalue mu_i mu_j mu_k, mu_i = (h/6)(4 + 2 cos(i pi h))."""
    th = np.arange(1, m + 1) * np.pi / (m + 1)
    t = 2.0 - 2.0 * np.cos(th)
    mu = (1.0 / (m + 1) / 6.0) * (4.0 + 2.0 * np.cos(th))
    c = min(m, int(np.ceil(nev ** (1.0 / 3.0))) + 6)
    lam = (t[:c, None, None] + t[None, :c, None] + t[None, None, :c]) / (
        mu[:c, None, None] * mu[None, :c, None] * mu[None, None, :c])
    prod = mu[:c, None, None] * mu[None, :c, None] * mu[None, None, :c]
    return float(prod.ravel()[np.argsort(lam.ravel(), kind="stable")[:nev]].min())


def matrix(case):
    if case["gen"] == "lap2d":
        return P.lap2d(case["m"])
    if case["gen"] == "lap3d_mass":
        return P.lap3d_mass(case["m"])
    return P.vardiff2d(case["m"], seed=case["gen_seed"])


def fingerprint(A):
    return np.array([A.shape[0], A.nnz, int(A.indptr.astype(np.int64).sum()),
                     int(A.indices.astype(np.int64).sum())], dtype=np.int64), float(A.data.sum())


def main(names):
    os.environ["MKL_NUM_THREADS"] = "1"
    if not O.ref_available():
        O.build_ref()
    for name in names:
        case = CASES[name]
        A, B = matrix(case)
        n, nev = A.shape[0], case["nev"]
        if B is not None:
            # every wanted eigenvalue is above 1, so "R" is ||r|| / (lambda ||x||)
            # and NS = R ||x|| / ||Bx|| <= R / min ||Bx|| / ||x||
            lam1 = float("nan")
            tol = float("%.3e" % (1e-8 * 0.95 * mass_floor(case["m"], nev)))
        else:
            lam1 = float(sla.eigsh(A.tocsc(), k=1, sigma=0.0, which="LM", return_eigenvectors=False)[0])
            tol = float("%.3e" % (1e-8 * 0.95 * lam1))
        flags = ["-gcge_ev_max_it", "2000", "-gcge_ev_tol", repr(tol), "-gcge_conv_type", "R",
                 "-gcge_print_para", "0"]
        X0 = P.random_x0(n, nev, case["x0_seed"])
        r = O.ref_solve(A, B, nev, X0=X0, flags=flags, timeout=7200)
        its, _ = O.parse_trace(r["stdout"])
        ints, dsum = fingerprint(A)
        np.savez_compressed(
            os.path.join(HERE, f"{name}_ref.npz"),
            gen=case["gen"], m=case["m"], gen_seed=case["gen_seed"], nev=nev, x0_seed=case["x0_seed"],
            lam1=lam1, ev_tol=tol, flags=np.array(flags), fp_ints=ints, fp_data_sum=dsum,
            eval=r["eval"], res=r["res"], num_iter=r["info"]["num_iter"], num_unlock=r["info"]["num_unlock"],
            seconds=r["info"].get("seconds", float("nan")),
            hist_eval=np.array([it["eval"] for it in its]), hist_res=np.array([it["res"] for it in its]),
            hist_unlock=np.array([it["num_unlock"] for it in its], dtype=np.int32))
        print(f"{name}: n={n} nev={nev} tol={tol} iters={r['info']['num_iter']} "
              f"unconv={r['info']['num_unlock']} lambda {r['eval'].min():.15e} .. {r['eval'].max():.15e}",
              flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))
