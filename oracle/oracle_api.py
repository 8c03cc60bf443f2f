"""Python access to the parity checker (test infrastructure only).

* ``solve(...)``      runs the C restatement (liboracle_gcge.so) in-process.
* ``ref_solve(...)``  runs the real reference binary oracle/_ref/ref_solve,
                      built from the reference's own sources by ``make ref``.
* ``parse_trace``     parses the per-iteration report both of them print
                      (GCGE_PrintIterationInfo format, print_eval = 1).
* kernel-level restatements: ``csr_matmultivec``, ``multivec_innerprod``,
  ``multivec_lincomb``, ``bcg``.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_gcge.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_BIN = os.path.join(REF_DIR, "ref_solve")

_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)


class OrcCSR(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("rowptr", _IP), ("col", _IP), ("val", _DP)]


class OrcParams(C.Structure):
    _fields_ = [
        ("nev", C.c_int), ("ev_max_it", C.c_int), ("block_size", C.c_int),
        ("given_init_evec", C.c_int), ("ev_tol", C.c_double),
        ("conv_type", C.c_char_p), ("orth_type", C.c_char_p),
        ("w_orth_type", C.c_char_p), ("p_orth_type", C.c_char_p),
        ("conv_omega_norm", C.c_double), ("orth_zero_tol", C.c_double),
        ("reorth_tol", C.c_double), ("scbgs_reorth_tol", C.c_double),
        ("criterion_tol", C.c_double), ("max_reorth_time", C.c_int),
        ("max_direct_orth_length", C.c_int), ("multi_tol_for_lock", C.c_double),
        ("cg_max_it", C.c_int), ("cg_rate", C.c_double), ("cg_type", C.c_int),
        ("if_use_bcg", C.c_int), ("dirichlet_boundary", C.c_int),
        ("opt_rr_eig_partly", C.c_int), ("seed", C.c_uint), ("w_orth_block_size", C.c_int),
        ("scbgs_wself_max_reorth_time", C.c_int),
    ]


class OrcInfo(C.Structure):
    _fields_ = [("num_iter", C.c_int), ("num_unlock", C.c_int), ("sum_res", C.c_double),
                ("bcg_calls", C.c_int), ("bcg_iters_total", C.c_int), ("spmv_count", C.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def build_ref():
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_default_params.argtypes = [C.POINTER(OrcParams)]
        L.orc_solve_file.argtypes = [C.POINTER(OrcCSR), C.POINTER(OrcCSR), C.POINTER(OrcParams),
                                     _DP, _DP, _DP, _DP, C.POINTER(OrcInfo), C.c_char_p]
        L.orc_solve_file.restype = C.c_int
        L.orc_csr_matmultivec.argtypes = [C.POINTER(OrcCSR), _DP, C.c_int, _DP, C.c_int, C.c_int]
        L.orc_multivec_innerprod.argtypes = [_DP, C.c_int, C.c_int, _DP, C.c_int, C.c_int, _DP, C.c_int]
        L.orc_multivec_lincomb.argtypes = [_DP, C.c_int, C.c_int, _DP, C.c_int, C.c_int, _DP,
                                           C.c_int, C.c_int]
        L.orc_bcg.argtypes = [C.POINTER(OrcCSR), _DP, _DP, C.c_int, C.c_int, C.c_int, C.c_double,
                              _IP]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


class _Keep:
    """Holds numpy arrays alive while a C struct points at them."""

    def __init__(self, M):
        self.rowptr = np.ascontiguousarray(M.indptr, dtype=np.int32)
        self.col = np.ascontiguousarray(M.indices, dtype=np.int32)
        self.val = np.ascontiguousarray(M.data, dtype=np.float64)
        self.s = OrcCSR(M.shape[0], self.col.size, _ip(self.rowptr), _ip(self.col), _dp(self.val))


def default_params(**kw):
    p = OrcParams()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        if isinstance(v, str):
            v = v.encode()
        setattr(p, k, v)
    return p


def solve(A, B, nev, X0=None, trace_path=None, **params):
    """Run the restatement. Returns dict(eval, evec, res, info)."""
    L = lib()
    n = A.shape[0]
    ka = _Keep(A)
    kb = _Keep(B) if B is not None else None
    if X0 is not None:
        params.setdefault("given_init_evec", 1)
        X0 = np.asfortranarray(X0, dtype=np.float64)
    p = default_params(nev=nev, **params)
    ev = np.zeros(nev)
    X = np.zeros((n, nev), order="F")
    res = np.zeros(nev)
    info = OrcInfo()
    rc = L.orc_solve_file(C.byref(ka.s), C.byref(kb.s) if kb else None, C.byref(p),
                          _dp(X0) if X0 is not None else None, _dp(ev), _dp(X), _dp(res),
                          C.byref(info), (trace_path or "").encode())
    if rc != 0:
        raise RuntimeError(f"orc_solve failed rc={rc}")
    return dict(eval=ev, evec=X, res=res,
                info=dict(num_iter=info.num_iter, num_unlock=info.num_unlock,
                          sum_res=info.sum_res, bcg_calls=info.bcg_calls,
                          bcg_iters_total=info.bcg_iters_total, spmv_count=info.spmv_count))


def write_problem(path, A, B, nev, X0=None):
    n = A.shape[0]
    with open(path, "wb") as f:
        f.write(b"GCGEPRB1")
        hdr = np.array([n, A.nnz, 1 if B is not None else 0, B.nnz if B is not None else 0,
                        nev, 1 if X0 is not None else 0], dtype=np.int32)
        f.write(hdr.tobytes())
        for M in ([A] if B is None else [A, B]):
            f.write(np.ascontiguousarray(M.indptr, dtype=np.int32).tobytes())
            f.write(np.ascontiguousarray(M.indices, dtype=np.int32).tobytes())
            f.write(np.ascontiguousarray(M.data, dtype=np.float64).tobytes())
        if X0 is not None:
            f.write(np.asfortranarray(X0, dtype=np.float64).tobytes(order="F"))


def read_output(path):
    with open(path, "rb") as f:
        assert f.read(8) == b"GCGEOUT1"
        n, nev, num_iter, num_unlock = np.frombuffer(f.read(16), dtype=np.int32)
        secs = np.frombuffer(f.read(8), dtype=np.float64)[0]
        ev = np.frombuffer(f.read(8 * nev), dtype=np.float64).copy()
        res = np.frombuffer(f.read(8 * nev), dtype=np.float64).copy()
        X = np.frombuffer(f.read(8 * n * nev), dtype=np.float64).reshape((nev, n)).T.copy(order="F")
    return dict(eval=ev, res=res, evec=X, seconds=float(secs),
                info=dict(num_iter=int(num_iter), num_unlock=int(num_unlock)))


def ref_available():
    return os.path.exists(REF_BIN)


def ref_solve(A, B, nev, X0=None, flags=(), seed=1, dirichlet=0, timeout=600, env=None):
    """Run the real reference on (A, B). Returns the output dict plus 'stdout'."""
    if not ref_available():
        raise RuntimeError("oracle/_ref/ref_solve is not built")
    with tempfile.TemporaryDirectory() as d:
        prob = os.path.join(d, "p.bin")
        out = os.path.join(d, "o.bin")
        write_problem(prob, A, B, nev, X0)
        cmd = [REF_BIN, prob, out, "-seed", str(seed), "-dirichlet", str(dirichlet),
               "-gcge_print_eval", "1"] + [str(x) for x in flags]
        e = dict(os.environ)
        e.setdefault("MKL_NUM_THREADS", "1")
        if env:
            e.update(env)
        cp = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=e)
        if cp.returncode != 0:
            raise RuntimeError(f"ref_solve failed: {cp.stderr[-2000:]}")
        r = read_output(out)
        r["stdout"] = cp.stdout
        return r


_NU = re.compile(r"^num_unlock\(\s*(\d+)\) =\s*(\d+); max_res\(\s*\d+\) = (\S+); min_res\(\s*\d+\) = (\S+);")
_EV = re.compile(r"^eval\((\d+),\s*(\d+)\) = (\S+) ; \S+\(\d+,\s*\d+\) =\s+(\S+) ;")
_SO = re.compile(r"^in GCGE_SubOrthonormalization")


def parse_trace(text):
    """List of per-iteration dicts: it, num_unlock, max_res, min_res, eval[], res[];
    plus the count of unconverged-reorthogonalization notices."""
    its = []
    notices = 0
    for line in text.splitlines():
        m = _NU.match(line)
        if m:
            its.append(dict(it=int(m.group(1)), num_unlock=int(m.group(2)),
                            max_res=float(m.group(3)), min_res=float(m.group(4)),
                            eval=[], res=[]))
            continue
        m = _EV.match(line)
        if m and its:
            its[-1]["eval"].append(float(m.group(3)))
            its[-1]["res"].append(float(m.group(4)))
            continue
        if _SO.match(line):
            notices += 1
    return its, notices


# ---- kernel-level restatements -------------------------------------------

def csr_matmultivec(A, X):
    L = lib()
    ka = _Keep(A)
    X = np.asfortranarray(X, dtype=np.float64)
    n, k = X.shape
    Y = np.zeros((n, k), order="F")
    L.orc_csr_matmultivec(C.byref(ka.s), _dp(X), n, _dp(Y), n, k)
    return Y


def multivec_innerprod(V, W):
    L = lib()
    V = np.asfortranarray(V, dtype=np.float64)
    W = np.asfortranarray(W, dtype=np.float64)
    n, p = V.shape
    q = W.shape[1]
    a = np.zeros((p, q), order="F")
    L.orc_multivec_innerprod(_dp(V), n, p, _dp(W), n, q, _dp(a), p)
    return a


def multivec_lincomb(X, a):
    L = lib()
    X = np.asfortranarray(X, dtype=np.float64)
    a = np.asfortranarray(a, dtype=np.float64)
    n, p = X.shape
    q = a.shape[1]
    Y = np.zeros((n, q), order="F")
    L.orc_multivec_lincomb(_dp(X), n, p, _dp(Y), n, q, _dp(a), a.shape[0], n)
    return Y


def bcg(A, RHS, W0, cg_max_it=20, cg_rate=1e-2):
    """GCGE_BCG restated: returns (W, final residual block, iterations)."""
    L = lib()
    ka = _Keep(A)
    R = np.array(RHS, dtype=np.float64, order="F", copy=True)
    W = np.array(W0, dtype=np.float64, order="F", copy=True)
    n, k = W.shape
    it = C.c_int(0)
    L.orc_bcg(C.byref(ka.s), _dp(R), _dp(W), n, k, cg_max_it, cg_rate, C.byref(it))
    return W, R, it.value
