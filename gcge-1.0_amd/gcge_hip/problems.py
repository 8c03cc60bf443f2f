"""Seeded generators for the GCGE benchmark problems.

Every generator returns ``(A, B)`` as ``scipy.sparse.csr_matrix`` with int32
indices and float64 values, rows sorted by column; ``B is None`` means the
identity.  Matrices are symmetric; A is positive definite.

Workloads (sizes are the defaults; tests use smaller ones):

* ``lap2d(128)``          2-D 5-point Dirichlet Laplacian, stencil [-1,-1,4,-1,-1]
                          on a 128x128 interior grid (n = 16384).
* ``lap3d_mass(96)``      3-D 7-point Dirichlet Laplacian on 96^3 with the
                          tensor-product P1 mass matrix (h^3/216)[1,4,1]^(x3).
* ``vardiff3d(160)``      cell-centred finite-volume diffusion on 160^3 cells,
                          kappa = exp(g), g a Gaussian field (sigma 1,
                          correlation length 8 cells), harmonic-mean face
                          transmissibilities, Dirichlet walls.
* ``chung_lu(2_000_000)`` graph Laplacian of a Chung-Lu random graph
                          (power-law exponent 2.5, mean degree 16) + 1e-2 I.
* ``vardiff2d(512)``      2-D finite-volume diffusion, x-faces scaled by 1.3,
                          theta = exp(g), g i.i.d. N(0, 0.25) (variance 0.25).

Closed-form spectra for the constant-coefficient problems are provided for
tests (``lap2d_eigs``, ``lap3d_mass_eigs``).  The 2-D Laplacian formula is the
one the reference computes in test/hypre/laplace.c:31-43.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _finish(M):
    M = sp.csr_matrix(M)
    M.sum_duplicates()
    M.sort_indices()
    M.indptr = M.indptr.astype(np.int32)
    M.indices = M.indices.astype(np.int32)
    M.data = M.data.astype(np.float64)
    return M


def _lap1d(m, diag=2.0):
    e = np.ones(m)
    return sp.diags([-e[:-1], diag * e, -e[:-1]], [-1, 0, 1], format="csr")


def lap2d(nx: int):
    """5-point Dirichlet Laplacian on an nx-by-nx interior grid (unscaled)."""
    T = _lap1d(nx)
    I = sp.identity(nx, format="csr")
    return _finish(sp.kron(I, T) + sp.kron(T, I)), None


def lap2d_eigs(nx: int, k: int):
    """Lowest k eigenvalues of lap2d(nx): 4 sin^2(i pi/(2(nx+1))) + 4 sin^2(j pi/(2(nx+1)))."""
    i = np.arange(1, nx + 1)
    s = 4.0 * np.sin(i * np.pi / (2.0 * (nx + 1))) ** 2
    ev = (s[:, None] + s[None, :]).ravel()
    ev.sort()
    return ev[:k]


def lap3d_mass(m: int):
    """3-D 7-point Laplacian (unscaled, [-1 x6, 6]) and P1 mass (h^3/216)[1,4,1]^(x3), h=1/(m+1)."""
    T = _lap1d(m)
    I = sp.identity(m, format="csr")
    A = sp.kron(sp.kron(I, I), T) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(T, I), I)
    e = np.ones(m)
    M1 = sp.diags([e[:-1], 4.0 * e, e[:-1]], [-1, 0, 1], format="csr")
    h = 1.0 / (m + 1)
    B = (h ** 3 / 216.0) * sp.kron(sp.kron(M1, M1), M1)
    return _finish(A), _finish(B)


def lap3d_mass_eigs(m: int, k: int):
    """Lowest k eigenvalues of lap3d_mass(m): (t_i+t_j+t_k)/(mu_i mu_j mu_k)."""
    th = np.arange(1, m + 1) * np.pi / (m + 1)
    t = 2.0 - 2.0 * np.cos(th)
    h = 1.0 / (m + 1)
    mu = (h / 6.0) * (4.0 + 2.0 * np.cos(th))
    # only the lowest indices matter: restrict to a cube that surely contains them
    c = min(m, int(np.ceil(k ** (1.0 / 3.0))) + 6)
    tt = t[:c]
    mm = mu[:c]
    lam = (tt[:, None, None] + tt[None, :, None] + tt[None, None, :]) / (
        mm[:, None, None] * mm[None, :, None] * mm[None, None, :])
    lam = np.sort(lam.ravel())
    return lam[:k]


def gaussian_field(shape, corr_len: float, sigma: float, seed: int):
    """Stationary Gaussian field: seeded white noise filtered by a Gaussian kernel
    of standard deviation corr_len (in cells) via FFT, rescaled to std sigma."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    freqs = [np.fft.fftfreq(s) for s in shape]
    k2 = np.zeros(shape)
    for ax, f in enumerate(freqs):
        sh = [1] * len(shape)
        sh[ax] = -1
        k2 = k2 + (f.reshape(sh) ** 2)
    filt = np.exp(-2.0 * (np.pi ** 2) * (corr_len ** 2) * k2)
    g = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    g -= g.mean()
    g *= sigma / g.std()
    return g


def fv_diffusion(kappa: np.ndarray, face_scale=None):
    """Cell-centred FV operator for -div(kappa grad u) on a unit-spacing grid with
    homogeneous Dirichlet walls.  Interior faces use the harmonic mean
    2 k_i k_j / (k_i + k_j); a wall face adds 2 k_i to the diagonal.
    face_scale[axis] multiplies all transmissibilities normal to that axis.
    Cells are numbered C-order (last axis fastest)."""
    shape = kappa.shape
    d = len(shape)
    n = kappa.size
    idx = np.arange(n).reshape(shape)
    diag = np.zeros(n)
    rows, cols, vals = [], [], []
    for ax in range(d):
        s = 1.0 if face_scale is None else float(face_scale[ax])
        lo = [slice(None)] * d
        hi = [slice(None)] * d
        lo[ax] = slice(0, shape[ax] - 1)
        hi[ax] = slice(1, shape[ax])
        k1 = kappa[tuple(lo)]
        k2 = kappa[tuple(hi)]
        t = s * 2.0 * k1 * k2 / (k1 + k2)
        i1 = idx[tuple(lo)].ravel()
        i2 = idx[tuple(hi)].ravel()
        t = t.ravel()
        np.add.at(diag, i1, t)
        np.add.at(diag, i2, t)
        rows += [i1, i2]
        cols += [i2, i1]
        vals += [-t, -t]
        for side in (0, shape[ax] - 1):
            sl = [slice(None)] * d
            sl[ax] = slice(side, side + 1)
            np.add.at(diag, idx[tuple(sl)].ravel(), s * 2.0 * kappa[tuple(sl)].ravel())
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(diag)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    return _finish(A)


def fv_diffusion_jacobian(kappa: np.ndarray, A, face_scale=None):
    """Jacobian of the stored values of A = fv_diffusion(kappa, face_scale)
    with respect to the cell coefficients, as CSR over cells: for cell c,
    jnz[jrowptr[c]:jrowptr[c+1]] are positions in A.data and jval the
    derivatives dA.data[pos] / dkappa_c.  An interior face t = 2 k1 k2/(k1+k2)
    has dt/dk1 = 2 k2^2/(k1+k2)^2; a wall face adds 2 k_i to the diagonal."""
    shape = kappa.shape
    d = len(shape)
    n = kappa.size
    idx = np.arange(n).reshape(shape)
    keys = np.repeat(np.arange(n, dtype=np.int64), np.diff(A.indptr)) * n + A.indices.astype(np.int64)

    def pos(r, c):
        return np.searchsorted(keys, r.astype(np.int64) * n + c.astype(np.int64))

    P, E, V = [], [], []
    for ax in range(d):
        s = 1.0 if face_scale is None else float(face_scale[ax])
        lo = [slice(None)] * d
        hi = [slice(None)] * d
        lo[ax] = slice(0, shape[ax] - 1)
        hi[ax] = slice(1, shape[ax])
        k1 = kappa[tuple(lo)].ravel()
        k2 = kappa[tuple(hi)].ravel()
        i1 = idx[tuple(lo)].ravel()
        i2 = idx[tuple(hi)].ravel()
        den = (k1 + k2) ** 2
        for cell, dt in ((i1, s * 2.0 * k2 * k2 / den), (i2, s * 2.0 * k1 * k1 / den)):
            for (r, c, sign) in ((i1, i2, -1.0), (i2, i1, -1.0), (i1, i1, 1.0), (i2, i2, 1.0)):
                P.append(cell)
                E.append(pos(r, c))
                V.append(sign * dt)
        for side in (0, shape[ax] - 1):
            sl = [slice(None)] * d
            sl[ax] = slice(side, side + 1)
            cells = idx[tuple(sl)].ravel()
            P.append(cells)
            E.append(pos(cells, cells))
            V.append(np.full(cells.size, s * 2.0))
    J = sp.csr_matrix((np.concatenate(V), (np.concatenate(P), np.concatenate(E))), shape=(n, A.nnz))
    J.sum_duplicates()
    J.sort_indices()
    return J.indptr.astype(np.int32), J.indices.astype(np.int32), J.data.astype(np.float64)


def vardiff3d(m: int = 160, sigma: float = 1.0, corr_len: float = 8.0, seed: int = 0):
    """3-D variable-coefficient diffusion on m^3 cells, kappa = exp(g)."""
    g = gaussian_field((m, m, m), corr_len, sigma, seed)
    return fv_diffusion(np.exp(g)), None


def vardiff2d_theta(m: int = 512, var: float = 0.25, seed: int = 0):
    """Per-cell coefficient theta = exp(g), g i.i.d. N(0, var)."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.standard_normal((m, m)) * np.sqrt(var))


def vardiff2d(m: int = 512, var: float = 0.25, seed: int = 0, anis: float = 1.3, theta=None):
    """2-D variable-coefficient diffusion, x-normal faces scaled by anis."""
    if theta is None:
        theta = vardiff2d_theta(m, var, seed)
    # axis 1 (fastest, x) faces scaled by anis
    return fv_diffusion(theta, face_scale=(1.0, anis)), None


def chung_lu(n: int = 2_000_000, beta: float = 2.5, mean_deg: float = 16.0, seed: int = 0,
             shift: float = 1e-2):
    """Graph Laplacian + shift*I of a Chung-Lu graph.  Expected degrees
    w_i ~ (i + i0)^(-1/(beta-1)) scaled to the requested mean; m = n*mean/2
    edges are drawn with endpoints sampled proportionally to w; self loops and
    duplicates are removed; unit edge weights.  Isolated vertices are kept
    (their eigenvalue is exactly `shift`), which makes the lowest spectrum
    highly degenerate."""
    rng = np.random.default_rng(seed)
    i0 = 10.0
    w = (np.arange(n) + i0) ** (-1.0 / (beta - 1.0))
    p = w / w.sum()
    m = int(n * mean_deg / 2)
    cdf = np.cumsum(p)
    u = rng.random(m)
    v = rng.random(m)
    a = np.minimum(np.searchsorted(cdf, u), n - 1).astype(np.int64)
    b = np.minimum(np.searchsorted(cdf, v), n - 1).astype(np.int64)
    keep = a != b
    a, b = a[keep], b[keep]
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    key = np.unique(lo * n + hi)
    lo = key // n
    hi = key % n
    W = sp.csr_matrix((np.ones(lo.size), (lo, hi)), shape=(n, n))
    W = W + W.T
    deg = np.asarray(W.sum(axis=1)).ravel()
    L = sp.diags(deg + shift) - W
    return _finish(L), None


def random_x0(n: int, k: int, seed: int):
    """Initial block X0 ~ U[-1, 1], n x k, Fortran (column-major) order."""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.uniform(-1.0, 1.0, size=(n, k)))
