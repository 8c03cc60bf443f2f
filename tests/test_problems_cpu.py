"""Host-side generators: the FV Jacobian that feeds the eigenvalue-gradient
kernel is pinned by central differences of the assembled matrix values."""
import numpy as np

from gcge_hip import problems as P


def test_fv_jacobian_matches_matrix_differences():
    m, anis = 9, 1.3
    theta = P.vardiff2d_theta(m, seed=3)
    A, _ = P.vardiff2d(m, theta=theta, anis=anis)
    jrp, jnz, jval = P.fv_diffusion_jacobian(theta, A, face_scale=(1.0, anis))
    assert jrp.size == m * m + 1 and jnz.max() < A.nnz
    rng = np.random.default_rng(4)
    for c in rng.choice(m * m, 6, replace=False):
        h = 1e-6
        tp, tm = theta.copy().ravel(), theta.copy().ravel()
        tp[c] += h
        tm[c] -= h
        Ap, _ = P.vardiff2d(m, theta=tp.reshape(m, m), anis=anis)
        Am, _ = P.vardiff2d(m, theta=tm.reshape(m, m), anis=anis)
        # same pattern at every theta, so the stored values line up
        assert np.array_equal(Ap.indices, A.indices) and np.array_equal(Am.indptr, A.indptr)
        fd = (Ap.data - Am.data) / (2 * h)
        dj = np.zeros(A.nnz)
        dj[jnz[jrp[c]:jrp[c + 1]]] = jval[jrp[c]:jrp[c + 1]]
        np.testing.assert_allclose(dj, fd, rtol=1e-6, atol=1e-8)


def test_fv_jacobian_3d_row_sums():
    """Every face derivative enters its two diagonal entries with + and its two
    off-diagonal entries with -, so each cell's derivative matrix has zero row
    sums except for the wall terms 2 s."""
    kappa = np.exp(np.random.default_rng(5).standard_normal((4, 5, 3)) * 0.5)
    A = P.fv_diffusion(kappa)
    jrp, jnz, jval = P.fv_diffusion_jacobian(kappa, A)
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    walls = np.zeros(kappa.shape)
    for ax in range(3):
        sl = [slice(None)] * 3
        for side in (0, kappa.shape[ax] - 1):
            sl[ax] = slice(side, side + 1)
            walls[tuple(sl)] += 2.0
    walls = walls.ravel()
    for c in range(A.shape[0]):
        e = jnz[jrp[c]:jrp[c + 1]]
        s = np.bincount(rows[e], weights=jval[jrp[c]:jrp[c + 1]], minlength=A.shape[0])
        expect = np.zeros(A.shape[0])
        expect[c] = walls[c]
        np.testing.assert_allclose(s, expect, atol=1e-12)
