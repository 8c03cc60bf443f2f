"""Benchmark: eigenpairs/s to 1e-8 relative residual, mixed precision vs FP64
GCGE, on one MI355X.

Workload (default): 3-D variable-coefficient diffusion on 160^3 cells
(n = 4,096,000, 28.5M nonzeros), kappa = exp(g) with g a seeded Gaussian
field (sigma 1, correlation length 8 cells), B = I, lowest nev = 100
eigenpairs, X0 ~ U[-1,1] (seeded), every pair converged to
||A x - lambda x|| / (|lambda| ||x||) <= 1e-8.

A step is one complete mixed-precision solve from X0 with the matrix and X0
already in HBM (default mode "mixed": an FP32 phase -- FP32 matrix values and
multivectors, FP32 SpMM / block CG / MFMA Gram and LinComb tiles, CholQR W
orthonormalization, FP64 dense Rayleigh-Ritz and residual checks -- then FP64
refinement sweeps of GCG from the FP32 eigenvectors with FP32 inner solves;
inner solves preconditioned with the diagonal of A (--bcg-precond) and capped
at --cg-max-it CG steps in both phases).  After the timed steps, FP64 solves
from the same X0 are timed for the comparison (every solver is set up
before its timed call and the timed calls run with the kernel profiler off;
one extra profiled solve of the same work gives the per-kernel rates): the drop-in with the
reference's defaults ("fp64": unpreconditioned block CG of at most 20 steps),
and FP64 with every algorithm setting of the mixed run -- CholQR, the same
preconditioner and step cap, the dsyevd Rayleigh-Ritz -- so that only the
precision differs
("fp64_same_settings").  The device kernels are
timed with HIP events on the backend stream over the timed region; the CPU
baseline is the reference built from its own sources (oracle/_ref/ref_solve)
on one core, on a bounded sample (see ``cpu_baseline``).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gcge-1.0_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_problem(args):
    from gcge_hip import problems as P
    w = args.workload
    if w == "vardiff3d":
        A, B = P.vardiff3d(args.size, sigma=1.0, corr_len=8.0, seed=args.seed)
    elif w == "lap2d":
        A, B = P.lap2d(args.size)
    elif w == "lap3d_mass":
        A, B = P.lap3d_mass(args.size)
    elif w == "vardiff2d":
        A, B = P.vardiff2d(args.size, seed=args.seed)
    elif w == "chung_lu":
        A, B = P.chung_lu(args.size, seed=args.seed)
    else:
        raise SystemExit(f"unknown workload {w}")
    return A, B


def kernel_delta(after, before):
    out = {}
    for k, v in after.items():
        b = before.get(k, dict(count=0, ms=0.0, bytes=0.0, flops=0.0))
        out[k] = {f: v[f] - b[f] for f in v}
    return out


# profiler class -> the kernel template whose instantiations rocprofv3 lists;
# the SpMM classes time the staged-row-block kernel, the direct column
# gather, or (matrices with scattered column indices) the row staging plus
# the row-gather kernel, as the matrix's layout says (HIP_MatGatherKind)
_ROWS = " / k_rows_from_cols + k_spmm_rows<{}, {}, ...>"
ROCPROF_KERNEL = {
    "spmm_f64": "k_spmm_halo<double, false, ...> / k_spmm<double, false, ...>" + _ROWS.format("double", "false"),
    "spmm_f32": "k_spmm_halo<float, false, ...> / k_spmm<float, false, ...>" + _ROWS.format("float", "false"),
    "spmm_dot_f64": "k_spmm_halo<double, true, ...> / k_spmm<double, true, ...>" + _ROWS.format("double", "true"),
    "spmm_dot_f32": "k_spmm_halo<float, true, ...> / k_spmm<float, true, ...>" + _ROWS.format("float", "true"),
    "axpby": "k_axpby<...>", "dots": "k_dots<...>",
    "cg_update_f64": "k_cg_resid<double> / k_pcg_resid<double>",
    "cg_update_f32": "k_cg_resid<float> / k_pcg_resid<float>",
    "cg_search_f64": "k_cg_search<double, ...>", "cg_search_f32": "k_cg_search<float, ...>",
    "gram_mfma_f64": "k_gram_mfma / k_gram_slab<double, ...>",
    "gram_mfma_f32": "k_gram_mfma / k_gram_slab<float, ...>",
    "gram_thin": "k_gram_thin<...>", "lincomb_mfma_f64": "k_lincomb_mfma<double, ...>",
    "lincomb_mfma_f32": "k_lincomb_mfma<float, ...>", "lincomb_thin": "k_lincomb_thin<...>",
    "residual": "k_residual<...> / k_rows_from_cols + k_residual_rows<...>", "copy": "k_copy<...>", "swap": "k_swap<...>",
}


def load_traffic(run, workload):
    """HBM bytes per launch measured with rocprofv3 --pmc for the kernel
    classes of one run ("mixed", or "fp64_same": FP64 with the mixed
    path's settings) of the workload they were
    profiled on (profiles/traffic.json); nothing for another workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(run, {}) if d.get("workload") == workload else {}
    except (OSError, ValueError, AttributeError):
        return {}


def roofline(stats, run, workload):
    """Dominant kernel class (largest device time) and its achieved rate:
    algorithmic bytes (or flops, for the MFMA tiles) per launch / average
    launch time, from HIP events on the backend stream."""
    if not stats:
        return None, None
    name = max(stats, key=lambda k: stats[k]["ms"])
    s = stats[name]
    if s["ms"] <= 0:
        return name, None
    cnt = max(1, s["count"])
    tr = load_traffic(run, workload).get(name)
    common = dict(kernel=ROCPROF_KERNEL.get(name, name), launches=s["count"],
                  avg_ms=round(s["ms"] / cnt, 5), bytes_per_launch=round(s["bytes"] / cnt),
                  traffic=(tr or {}).get("bytes_per_launch"))
    if "mfma" in name:
        peak = FP64_PEAK_TFLOPS if name.endswith("f64") else FP32_PEAK_TFLOPS
        ach = s["flops"] / (s["ms"] * 1e-3) / 1e12
        return name, dict(bound="mfma", achieved=round(ach, 3), peak=peak, unit="TFLOP/s",
                          frac=round(ach / peak, 4), flops_per_launch=round(s["flops"] / cnt), **common)
    ach = s["bytes"] / (s["ms"] * 1e-3) / 1e9
    return name, dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(ach / HBM_PEAK_GBS, 4), **common)


def cpu_baseline(args, n_full, nnz_full, fp64_iters, nev):
    """Reference (oracle/_ref/ref_solve, one core) on a bounded sample: a few
    outer iterations on the same generator at a reduced size with the same
    nev; the per-iteration time is scaled linearly in the nonzero count to the
    full problem and multiplied by the GPU FP64 iteration count."""
    import oracle_api as O
    from gcge_hip import problems as P
    # sample sizes: a few seconds of reference work per iteration, beyond the
    # host's caches (the 2-D Laplacian is small enough to run at full size)
    m = args.cpu_sample_size or {"vardiff3d": 48, "lap3d_mass": 48, "lap2d": args.size, "vardiff2d": 256,
                                 "chung_lu": 20000}.get(args.workload, 48)
    its = args.cpu_sample_iters
    if args.workload == "vardiff3d":
        A, B = P.vardiff3d(m, seed=args.seed)
    elif args.workload == "lap2d":
        A, B = P.lap2d(m)
    elif args.workload == "lap3d_mass":
        A, B = P.lap3d_mass(m)
    elif args.workload == "vardiff2d":
        A, B = P.vardiff2d(m, seed=args.seed)
    else:
        A, B = P.chung_lu(m, seed=args.seed)
    X0 = P.random_x0(A.shape[0], nev, args.seed + 1)
    env = {"MKL_NUM_THREADS": "1", "OMP_NUM_THREADS": "1"}
    kind = "reference"
    t0 = time.perf_counter()
    if O.ref_available():
        r = O.ref_solve(A, B, nev, X0=X0, flags=["-gcge_ev_max_it", str(its), "-gcge_print_para", "0",
                                                   "-gcge_print_result", "0"], env=env, timeout=600)
        secs, iters = r["seconds"], r["info"]["num_iter"]
    else:
        kind = "port"
        t1 = time.perf_counter()
        o = O.solve(A, B, nev, X0=X0, ev_max_it=its)
        secs, iters = time.perf_counter() - t1, o["info"]["num_iter"]
    per_iter = secs / max(1, iters)
    scale = (nnz_full + 2.0 * n_full) / (A.nnz + 2.0 * A.shape[0])
    projected = per_iter * scale * max(1, fp64_iters)
    log(f"cpu baseline: {kind} {iters} iterations at n={A.shape[0]} in {secs:.2f}s "
        f"({time.perf_counter() - t0:.1f}s wall), projected {projected:.1f}s for the full solve")
    model, nproc = host_cpu()
    return dict(value=nev / projected, unit="eigenpairs/s", cores=1, kind=kind, host_cpu=model, host_nproc=nproc,
                sample=(f"{iters} outer iterations of the reference solve (1 core, MKL 1 thread) on "
                        f"{args.workload} size {m} (n={A.shape[0]}, nnz={A.nnz}) with nev={nev}; "
                        f"time per iteration scaled by (nnz+2n) to the full problem and multiplied by "
                        f"the GPU FP64 iteration count {fp64_iters}"),
                seconds_measured=round(secs, 3), projected_seconds=round(projected, 2))


def copy_probe():
    from gcge_hip import Matrix, MultiVec, Ops, capi
    from gcge_hip import problems as P
    A, _ = P.lap2d(2024)                  # n = 4,096,576 rows
    mA = Matrix(A)
    k = 16
    x = MultiVec(mA, k)
    y = MultiVec(mA, k)
    ops = Ops()
    for _ in range(3):
        ops.copy(x, y)
    capi.load().GCGE_HIP_Synchronize()
    n = A.shape[0]
    print(json.dumps({"copy_probe": {"kernel": "k_copy<double, double>", "rows": n, "columns": k,
                                     "read_bytes": 8 * n * k, "write_bytes": 8 * n * k, "launches": 3}}))


def spmm_probe(args):
    """Three launches each of the FP64 and the FP32 SpMM with fused dots over
    36 and 100 columns on the workload's matrix, for profiler runs
    (rocprofv3 --kernel-trace / --pmc); prints the HIP-event rate of each."""
    from gcge_hip import Matrix, MultiVec, Ops, capi
    L = capi.load()
    A, _ = make_problem(args)
    mA = Matrix(A, order=row_order(args))
    ops = Ops()
    out = []
    for prec, w in ((capi.GCGE_HIP_FP64, 8), (capi.GCGE_HIP_FP32, 4)):
        for k in (36, 100):
            V = MultiVec(mA, k, prec)
            W = MultiVec(mA, k, prec)
            ops.axpby(0.0, V, 0.0, V)
            ops.matdot_dot(mA, V, W)
            L.GCGE_HIP_Synchronize()
            L.GCGE_HIP_ResetKernelStats()
            L.GCGE_HIP_SetProfiling(1)
            for _ in range(3):
                ops.matdot_dot(mA, V, W)
            L.GCGE_HIP_Synchronize()
            L.GCGE_HIP_SetProfiling(0)
            st = capi.kernel_stats()
            cls = "spmm_dot_f64" if prec == capi.GCGE_HIP_FP64 else "spmm_dot_f32"
            s = st[cls]
            alg = (w + 4) * A.nnz + 4 * (A.shape[0] + 1) + 2 * k * A.shape[0] * w
            ms = s["ms"] / max(1, s["count"])
            out.append(dict(kernel=cls, columns=k, avg_ms=round(ms, 4), alg_bytes=alg,
                            gbs=round(alg / (ms * 1e-3) / 1e9, 1)))
            log(f"spmm probe {cls} k={k}: {ms:.4f} ms, {alg / (ms * 1e-3) / 1e9:.0f} GB/s")
            V.close()
            W.close()
    print(json.dumps({"spmm_probe": out, "workload": f"{args.workload} {args.size}"}))


def row_order(args):
    """Row order of the device matrices: "locality" (HIP_MatLocalityOrder:
    the grids come back in 8x8x8 / 16x16 tiles) for the grid workloads,
    None (the caller's order) for the random graph, whose scattered columns
    take the row-gather SpMM, or as --order says."""
    if args.order == "natural":
        return None
    if args.order == "locality":
        return "locality"
    return None if args.workload == "chung_lu" else "locality"


def spmm_kind(mA):
    """The SpMM layout the library chose for the matrix (HIP_MatGatherKind)."""
    from gcge_hip import capi
    return {0: "columns (k_spmm)", 1: "rows (k_spmm_rows)", 2: "staged row blocks (k_spmm_halo)"}.get(
        capi.load().HIP_MatGatherKind(mA.ptr), "?")


def host_cpu():
    """CPU model name and the cores this process may use (the box's share)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        nproc = os.cpu_count()
    return model, nproc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="vardiff3d")
    ap.add_argument("--size", type=int, default=160)
    ap.add_argument("--nev", type=int, default=100)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--max-it", type=int, default=2000)
    ap.add_argument("--fp32-tol", type=float, default=8e-3)
    ap.add_argument("--mode", default="mixed", choices=["mixed", "fp32+refine", "fp64-cg32", "fp64"],
                    help="mixed: FP32 phase (CholQR) then FP64 refinement with FP32 inner solves; "
                         "fp32+refine: the same phases with the reference's orthonormalization and FP64 inner "
                         "solves; fp64-cg32: FP64 with FP32 inner block-CG solves; fp64: the FP64 drop-in "
                         "itself as the timed step (for profiling it; use with --no-fp64 --no-fp64-same)")
    ap.add_argument("--verbose", action="store_true",
                    help="print GCGE's per-iteration convergence line (stdout, before the JSON line)")
    ap.add_argument("--stagnation-window", type=int, default=2)
    ap.add_argument("--stagnation-ratio", type=float, default=0.5)
    ap.add_argument("--stagnation-start", type=float, default=1e-3)
    ap.add_argument("--order", default="auto", choices=["auto", "locality", "natural"],
                    help="row order of the device matrices (auto: locality for the grid workloads)")
    ap.add_argument("--rr-eig-partly", type=int, default=1, choices=[0, 1],
                    help="soft locking (opt_rr_eig_partly) in the mixed and same-settings FP64 solves")
    ap.add_argument("--leg", default="all", choices=["all", "mixed", "fp64", "fp64-same"],
                    help="run one leg alone (mixed; FP64 with the reference's defaults; FP64 with the mixed "
                         "path's settings), for profiler runs; no JSON bench line")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 solve (reference defaults)")
    ap.add_argument("--cg-max-it", type=int, default=30,
                    help="CG step cap of the mixed path's inner solves (the FP64 drop-in keeps the reference's 20)")
    ap.add_argument("--bcg-precond", type=int, default=1, choices=[0, 1],
                    help="1: the mixed path's inner solves are preconditioned with the diagonal of A (the FP64 "
                         "drop-in keeps the reference's unpreconditioned block CG)")
    ap.add_argument("--no-fp64-same", action="store_true",
                    help="skip the FP64 solve with the mixed path's algorithm settings (CholQR W "
                         "orthonormalization, the same inner CG step cap)")
    ap.add_argument("--no-profiled-step", action="store_true",
                    help="skip the extra profiled solve that gives the per-kernel rates")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample-size", type=int, default=0,
                    help="generator size of the CPU sample (0: 48 for the 3-D grids, 256 for vardiff2d, "
                         "the full size for lap2d, 20000 vertices for chung_lu)")
    ap.add_argument("--cpu-sample-iters", type=int, default=3)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run k_copy<double, double> over 16 columns of n = 4,096,000 rows three times "
                         "(1 GiB moved per launch), the reference of known size for rocprofv3 --pmc byte counters")
    ap.add_argument("--spmm-probe", action="store_true",
                    help="only time the FP64 and FP32 SpMM (fused dots, 36 and 100 columns) on the workload's "
                         "matrix, three launches each, for profiler runs")
    args = ap.parse_args()

    if args.copy_probe:
        copy_probe()
        return
    if args.spmm_probe:
        spmm_probe(args)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        log("bench.py: only --gpus 1 is supported (one process drives one GPU)")
        sys.exit(2)

    from gcge_hip import Matrix, MultiVec, Solver, capi
    from gcge_hip import problems as P
    L = capi.load()

    t = time.perf_counter()
    A, B = make_problem(args)
    n, nnz, nev = A.shape[0], A.nnz, args.nev
    X0 = P.random_x0(n, nev, args.seed + 1)
    log(f"problem {args.workload} size {args.size}: n={n} nnz={nnz} nev={nev} ({time.perf_counter() - t:.1f}s)")
    t = time.perf_counter()
    order = row_order(args)
    mA = Matrix(A, order=order)
    mB = Matrix(B, order=mA.order()) if B is not None else None
    X0 = MultiVec(mA, nev, data=X0)  # resident in HBM before any timed solve
    L.GCGE_HIP_Synchronize()
    log(f"upload of the matrix and X0 over PCIe: {time.perf_counter() - t:.2f}s")

    common = dict(conv_type="NS", ev_tol=args.tol, ev_max_it=args.max_it, print_result=1 if args.verbose else 0)
    # settings of the mixed path that the same-settings FP64 solve shares
    # (the FP64 drop-in keeps the reference's defaults)
    tuned = dict(opt_rr_eig_partly=args.rr_eig_partly)
    mixed = dict(fp32_tol=args.fp32_tol, fp32_max_it=args.max_it, fp64_max_it=args.max_it,
                 stagnation_window=args.stagnation_window, stagnation_ratio=args.stagnation_ratio,
                 stagnation_start=args.stagnation_start, cg_max_it=args.cg_max_it, bcg_precond=args.bcg_precond)

    def ns_of(solver, res):
        """NS residual of every pair, recomputed in FP64 on the device, and
        the strict convergence verdict: no unconverged pair reported, every
        residual finite and within the tolerance."""
        rn, xn, bn = solver.residual_norms()
        ev = np.asarray(res["eval"])
        ns = rn / (np.abs(ev) * bn)
        ok = res["num_unlock"] == 0 and bool(np.all(np.isfinite(ns))) and float(ns.max()) <= args.tol * (1 + 1e-12)
        return ns, ok

    def timed(solver, what):
        L.GCGE_HIP_Synchronize()
        t0 = time.perf_counter()
        r = solver.solve(X0, keep_evec=False)
        L.GCGE_HIP_Synchronize()
        dt = time.perf_counter() - t0
        log(f"{what}: {dt:.3f}s  {r['num_iter']} it, unconverged {r['num_unlock']} {r['info']}")
        return r, dt

    def profiled(solver, what):
        """One more solve of the same work with the per-launch HIP-event
        profiler on: kernel classes, launch counts and device time."""
        L.GCGE_HIP_Synchronize()
        L.GCGE_HIP_SetProfiling(1)
        before = capi.kernel_stats()
        t0 = time.perf_counter()
        r = solver.solve(X0, keep_evec=False)
        L.GCGE_HIP_Synchronize()
        dt = time.perf_counter() - t0
        stats = kernel_delta(capi.kernel_stats(), before)
        L.GCGE_HIP_SetProfiling(0)
        kms = sum(v["ms"] for v in stats.values())
        log(f"{what} (profiled): {dt:.3f}s, kernels {kms / 1e3:.3f}s")
        return r, dt, stats, kms

    if args.leg != "all":
        # one leg alone, for profiler runs (rocprofv3 --kernel-trace / --pmc):
        # the solver of that leg, args.warmup + args.steps solves from X0
        kw = dict(precision="mixed", mixed=mixed, **tuned) if args.leg == "mixed" else dict(precision="fp64")
        if args.leg == "fp64-same":
            kw.update(w_orth_type="cholqr", cg_max_it=args.cg_max_it, bcg_precond=args.bcg_precond, rr_eig_full=1,
                      **tuned)
        S = Solver(mA, mB, nev, **kw, **common)
        out = []
        for i in range(args.warmup + args.steps):
            r, dt = timed(S, f"{args.leg} solve {i}")
            out.append(dict(seconds=round(dt, 4), iters=r["num_iter"], converged=bool(ns_of(S, r)[1])))
        S.close()
        print(json.dumps({"leg": args.leg, "workload": f"{args.workload} {args.size}", "nev": nev, "solves": out}))
        return

    # ---- the timed steps: one solver set up once (workspace allocated
    # outside the timed region), solves from the same X0, profiler off
    S = Solver(mA, mB, nev, precision=args.mode, mixed=mixed, **(tuned if args.mode != "fp64" else {}), **common)
    warm = []
    for i in range(args.warmup):
        warm.append(timed(S, f"warmup {i}")[1])
    L.GCGE_HIP_Synchronize()
    t0 = time.perf_counter()
    step_s = []
    last = None
    for i in range(args.steps):
        ts = time.perf_counter()
        last = S.solve(X0, keep_evec=False)
        L.GCGE_HIP_Synchronize()
        step_s.append(time.perf_counter() - ts)
        log(f"step {i}: {step_s[-1]:.3f}s  {last['num_iter']} it {last['info']}")
    L.GCGE_HIP_Synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = nev / (ms_per_step * 1e-3)
    ns, converged = ns_of(S, last)
    mixed_eval = np.asarray(last["eval"])
    prof = None
    if not args.no_profiled_step:
        rp, dtp, stats, kms = profiled(S, "profiled step")
        prof = dict(seconds=dtp, stats=stats, kernel_s=kms / 1e3, iters=rp["num_iter"])
    S.close()

    fp64 = None
    if not args.no_fp64:
        # the FP64 drop-in (reference defaults), set up outside the timed call
        S64 = Solver(mA, mB, nev, precision="fp64", **common)
        r64, t64 = timed(S64, "fp64 (reference defaults)")
        ns64, ok64 = ns_of(S64, r64)
        fp64 = dict(seconds=t64, iters=r64["num_iter"], converged=ok64, eval=np.asarray(r64["eval"]),
                    max_ns=float(np.max(ns64)))
        S64.close()
    fp64s = None
    if not args.no_fp64_same and args.mode == "mixed":
        # FP64 with every algorithm setting of the mixed run: timed like a step
        S64s = Solver(mA, mB, nev, precision="fp64", w_orth_type="cholqr", cg_max_it=args.cg_max_it,
                      bcg_precond=args.bcg_precond, rr_eig_full=1, **tuned, **common)
        # one untimed solve first, as the steps have their warm-up (the first
        # solve of a new solver grows scratch buffers and pages memory in)
        timed(S64s, "fp64 with the mixed path's settings (warm-up)")
        r64s, t64s = timed(S64s, "fp64 with the mixed path's settings")
        ns64s, ok64s = ns_of(S64s, r64s)
        fp64s = dict(seconds=t64s, iters=r64s["num_iter"], converged=ok64s, eval=np.asarray(r64s["eval"]),
                     max_ns=float(np.max(ns64s)))
        if not args.no_profiled_step:
            _, dtp, st64, kms64 = profiled(S64s, "fp64 same settings")
            fp64s.update(stats=st64, profiled_seconds=dtp, kernel_s=kms64 / 1e3)
        S64s.close()

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args, n, nnz + (B.nnz if B is not None else 0),
                               fp64["iters"] if fp64 else last["num_iter"], nev)
        except Exception as e:  # the baseline must not hide the GPU result
            log(f"cpu baseline failed: {e}")

    wl = f"{args.workload} {args.size}"
    stats = prof["stats"] if prof else {}
    dom, roof = roofline(stats, "fp64" if args.mode == "fp64" else "mixed", wl)
    line = {
        "metric": "Eigenpairs/sec to 1e-8 rel. residual, mixed vs FP64 GCGE, 1 MI355X",
        "value": round(value, 4),
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.mode == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} size {args.size} (n={n}, nnz={nnz}), nev={nev}, B=I"
                               if B is None else f"{args.workload} size {args.size} (n={n}, nnz={nnz}+{B.nnz}), nev={nev}",
                   "n": n, "nnz": nnz, "nev": nev, "tol": args.tol,
                   "criterion": "||Ax-lambda Bx||/(|lambda| ||Bx||)", "x0": "U[-1,1] seeded",
                   "mode": args.mode, "fp32_tol": args.fp32_tol, "cg_max_it": args.cg_max_it,
                   "bcg_precond": "jacobi" if args.bcg_precond else "none",
                   "rr_eig": "dsyevd" if args.mode == "mixed" else "dsyevx",
                   "soft_locking": bool(args.rr_eig_partly) if args.mode != "fp64" else False,
                   "stagnation": [args.stagnation_window, args.stagnation_ratio, args.stagnation_start],
                   "row_order": "locality (HIP_MatLocalityOrder)" if order else "natural",
                   "spmm": spmm_kind(mA)},
        "converged": bool(converged),
        "max_ns_residual": float(np.max(ns)),
        "step_seconds": [round(x, 3) for x in step_s],
        "warmup_seconds": [round(x, 3) for x in warm],
        "timing": "timed steps with the kernel profiler off; per-kernel rates from one extra profiled step",
        "mixed": dict(mode=args.mode, iters=last["num_iter"], bcg_sweeps=last["bcg_iters_total"],
                      **{k: (round(v, 3) if isinstance(v, float) else v) for k, v in last["info"].items()}),
        "roofline_kernel": dom,
        "roofline": roof,
        "kernels_mixed": {k: dict(count=v["count"], ms=round(v["ms"], 2),
                                  gbs=round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None)
                          for k, v in stats.items() if v["count"]},
        "cpu_baseline": cpu,
    }
    if prof:
        line["profiled_step"] = dict(seconds=round(prof["seconds"], 3), kernel_seconds=round(prof["kernel_s"], 3),
                                     outside_kernels_seconds=round(prof["seconds"] - prof["kernel_s"], 3))
    if fp64:
        line["fp64"] = dict(eigenpairs_per_s=round(nev / fp64["seconds"], 4), seconds=round(fp64["seconds"], 3),
                            iters=fp64["iters"], converged=fp64["converged"], max_ns_residual=fp64["max_ns"])
        line["mixed_vs_fp64"] = round(fp64["seconds"] / (ms_per_step * 1e-3), 3)
    if fp64s:
        # the precision-only comparison: FP64 with every algorithm setting of
        # the mixed path (CholQR, inner CG step cap and preconditioner, dense
        # Rayleigh-Ritz solver, soft locking), timed the same way as the steps
        line["fp64_same_settings"] = dict(eigenpairs_per_s=round(nev / fp64s["seconds"], 4),
                                          seconds=round(fp64s["seconds"], 3), iters=fp64s["iters"],
                                          converged=fp64s["converged"], max_ns_residual=fp64s["max_ns"],
                                          w_orth_type="cholqr", cg_max_it=args.cg_max_it,
                                          bcg_precond=args.bcg_precond, rr_eig="dsyevd",
                                          soft_locking=bool(args.rr_eig_partly))
        if "stats" in fp64s:
            d64, roof64 = roofline(fp64s["stats"], "fp64_same", wl)
            line["fp64_same_settings"].update(
                roofline_kernel=d64, roofline=roof64,
                kernels={k: dict(count=v["count"], ms=round(v["ms"], 2),
                                 gbs=round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None)
                         for k, v in fp64s["stats"].items() if v["count"]})
        line["mixed_vs_fp64_same_settings"] = round(fp64s["seconds"] / (ms_per_step * 1e-3), 3)
    ref = fp64s or fp64
    if ref:
        # the error of every mixed eigenvalue against the FP64 solve
        rel = np.abs(mixed_eval - ref["eval"]) / np.abs(ref["eval"])
        line["max_rel_eig_diff_mixed_vs_fp64"] = float(rel.max())
        line["rel_eig_diff_per_pair_vs_fp64"] = [float("%.2e" % x) for x in rel]
        line["rel_eig_diff_reference"] = "fp64_same_settings" if fp64s else "fp64"
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
