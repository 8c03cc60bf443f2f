"""Profiling probe for the column-gather SpMM: three launches each of the FP64
and the FP32 SpMM with fused dots, 36 columns, on the 160^3 variable-
coefficient diffusion operator (n = 4.1M, nnz = 28.5M).  Run it under a
profiler, e.g. rocprofv3 --kernel-trace --stats --pmc ... -- python3
gcge-1.0_amd/gcge_hip/spmm_probe.py; it prints one line per phase."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from gcge_hip import Matrix, MultiVec, Ops, capi  # noqa: E402
from gcge_hip import problems as P  # noqa: E402


def main(k=36, reps=3):
    L = capi.load()
    ops = Ops()
    A, _ = P.vardiff3d(160, seed=0)
    mA = Matrix(A)
    print("matrix ready", flush=True)
    for prec in (capi.GCGE_HIP_FP64, capi.GCGE_HIP_FP32):
        L.HIP_MatSetPrecision(mA.ptr, prec)
        V = MultiVec(mA, k, prec)
        W = MultiVec(mA, k, prec)
        ops.axpby(0.0, V, 0.0, V)
        for _ in range(reps):
            ops.matdot_dot(mA, V, W, k)
        L.GCGE_HIP_Synchronize()
        print("done", "fp64" if prec == capi.GCGE_HIP_FP64 else "fp32", flush=True)
        del V, W
    mA.close()


if __name__ == "__main__":
    main()
