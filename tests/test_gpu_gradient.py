"""Eigenvalue gradients through the adjoint kernel (GCGE_HIP_EigGradient).

loss(theta) = sum of the lowest 16 eigenvalues of the 2-D FV diffusion
operator with per-cell coefficient theta = exp(g), g i.i.d. N(0, 0.25), and
x-normal faces scaled by 1.3 so that the lowest spectrum is simple.  For
unit-norm eigenvectors d(lambda_i)/d(theta_c) = x_i^T (dA/dtheta_c) x_i (B = I).

- 32 x 32 cells: the device gradient of every cell against the same formula
  with eigenvectors from dense LAPACK (scipy.linalg.eigh); FP64 central
  differences of the dense loss with step 1e-6 along random directions, and
  for single cells with step 1e-4 (a single-cell derivative is ~1e-3, so at
  step 1e-6 the loss difference is below the eigensolver's rounding
  resolution of about 16 eps ||A||).
- 512 x 512 cells: a torch.autograd.Function whose forward runs the FP64
  device solver and whose backward runs the adjoint kernel; its gradient is
  checked against FP64 central differences (step 1e-6) along random
  directions, each loss evaluated by the device solver.
Each check asserts 1e-6 relative agreement.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from gcge_hip import Matrix, eig_gradient, solve
from gcge_hip import problems as P

pytestmark = pytest.mark.gpu

ANIS = 1.3
NEV = 16


def loss_dense(m, theta):
    A, _ = P.vardiff2d(m, theta=theta, anis=ANIS)
    w = sla.eigh(A.toarray(), eigvals_only=True, subset_by_index=[0, NEV - 1])
    return float(np.sum(w))


def assert_simple(ev, rel=1e-6):
    gaps = np.diff(np.sort(ev)) / np.abs(ev[1:])
    assert gaps.min() > rel, f"lowest spectrum is (nearly) degenerate: min relative gap {gaps.min():.2e}"


def test_gradient_32x32_dense_and_fd():
    m = 32
    theta = P.vardiff2d_theta(m, var=0.25, seed=11)
    A, _ = P.vardiff2d(m, theta=theta, anis=ANIS)
    jrp, jnz, jval = P.fv_diffusion_jacobian(theta, A, face_scale=(1.0, ANIS))
    w, V = sla.eigh(A.toarray(), subset_by_index=[0, NEV])
    assert_simple(w)
    r = solve(A, None, NEV, X0=P.random_x0(m * m, NEV, 12), conv_type="NS", ev_tol=1e-11, ev_max_it=2000)
    assert r["num_unlock"] == 0
    np.testing.assert_allclose(r["eval"], w[:NEV], rtol=1e-10)
    mA = Matrix(A)
    g_dev = eig_gradient(mA, r["evec"], np.ones(NEV), m * m, jrp, jnz, jval)
    mA.close()
    rows = np.repeat(np.arange(m * m), np.diff(A.indptr))
    ge = np.zeros(A.nnz)
    for i in range(NEV):
        ge += V[rows, i] * V[A.indices, i]
    g_dense = np.array([np.dot(jval[jrp[c]:jrp[c + 1]], ge[jnz[jrp[c]:jrp[c + 1]]]) for c in range(m * m)])
    np.testing.assert_allclose(g_dev, g_dense, rtol=1e-6, atol=1e-9 * np.abs(g_dense).max())
    rng = np.random.default_rng(13)
    h = 1e-6
    for _ in range(3):
        v = rng.uniform(0.0, 1.0, m * m)
        fd = (loss_dense(m, theta + h * v.reshape(m, m)) - loss_dense(m, theta - h * v.reshape(m, m))) / (2 * h)
        ad = float(np.dot(g_dev, v))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    h = 1e-4
    for c in rng.choice(m * m, 4, replace=False):
        tp, tm = theta.ravel().copy(), theta.ravel().copy()
        tp[c] += h
        tm[c] -= h
        fd = (loss_dense(m, tp.reshape(m, m)) - loss_dense(m, tm.reshape(m, m))) / (2 * h)
        assert abs(g_dev[c] - fd) <= 1e-6 * abs(fd), (c, g_dev[c], fd)


def make_eigsum(m, tol):
    import torch

    class EigSum(torch.autograd.Function):
        """loss = sum of the lowest NEV eigenvalues; backward = adjoint kernel."""

        @staticmethod
        def forward(ctx, theta):
            th = theta.detach().cpu().numpy().reshape(m, m)
            A, _ = P.vardiff2d(m, theta=th, anis=ANIS)
            r = solve(A, None, NEV + 1, X0=P.random_x0(m * m, NEV + 1, 21), conv_type="NS", ev_tol=tol,
                      ev_max_it=4000)
            assert r["num_unlock"] == 0
            ctx.th, ctx.A, ctx.X, ctx.ev = th, A, r["evec"][:, :NEV], r["eval"]
            return torch.tensor(float(np.sum(r["eval"][:NEV])), dtype=torch.float64)

        @staticmethod
        def backward(ctx, g):
            jrp, jnz, jval = P.fv_diffusion_jacobian(ctx.th, ctx.A, face_scale=(1.0, ANIS))
            mA = Matrix(ctx.A)
            grad = eig_gradient(mA, ctx.X, np.full(NEV, float(g)), m * m, jrp, jnz, jval)
            mA.close()
            return torch.from_numpy(grad)

    return EigSum


def loss_device(m, theta, tol, X0):
    A, _ = P.vardiff2d(m, theta=theta, anis=ANIS)
    r = solve(A, None, NEV + 1, X0=X0, conv_type="NS", ev_tol=tol, ev_max_it=4000)
    assert r["num_unlock"] == 0
    return float(np.sum(r["eval"][:NEV]))


@pytest.mark.timeout(900)
def test_gradient_512x512_autograd_against_fd():
    import torch
    m, tol = 512, 1e-10
    theta0 = P.vardiff2d_theta(m, var=0.25, seed=0)
    theta = torch.tensor(theta0.ravel(), dtype=torch.float64, requires_grad=True)
    EigSum = make_eigsum(m, tol)
    loss = EigSum.apply(theta)
    loss.backward()
    grad = theta.grad.numpy()
    assert np.all(np.isfinite(grad))
    # non-degenerate lowest spectrum (NEV + 1 values from the forward solve)
    A, _ = P.vardiff2d(m, theta=theta0, anis=ANIS)
    r = solve(A, None, NEV + 1, X0=P.random_x0(m * m, NEV + 1, 21), conv_type="NS", ev_tol=tol, ev_max_it=4000,
              keep_evec=True)
    assert_simple(r["eval"])
    h = 1e-6
    rng = np.random.default_rng(22)
    for _ in range(2):
        v = rng.uniform(0.0, 1.0, m * m)
        Lp = loss_device(m, (theta0.ravel() + h * v).reshape(m, m), tol, r["evec"])
        Lm = loss_device(m, (theta0.ravel() - h * v).reshape(m, m), tol, r["evec"])
        fd = (Lp - Lm) / (2 * h)
        ad = float(np.dot(grad, v))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


def laplace_plus_mass(m, kappa, rho):
    """A(kappa) = the FV diffusion operator (x-faces x1.3) and B(rho) = I +
    0.1 L(rho), L(rho) the FV operator with coefficients rho: both depend on
    cell parameters, B through off-diagonal entries too.  Returns A, B and
    the Jacobians of their stored values with respect to theta = [kappa; rho]
    (2 m^2 parameters): A's rows for kappa, B's for rho."""
    n = m * m
    A, _ = P.vardiff2d(m, theta=kappa, anis=ANIS)
    L = P.fv_diffusion(rho)
    B = L.copy()
    B.data = 0.1 * B.data
    diag = np.repeat(np.arange(n), np.diff(B.indptr)) == B.indices
    B.data[diag] += 1.0
    ja = P.fv_diffusion_jacobian(kappa, A, face_scale=(1.0, ANIS))
    jl = P.fv_diffusion_jacobian(rho, L)
    # stack to 2n parameter rows: A depends on the first n, B on the last n
    ja2 = (np.concatenate([ja[0], np.full(n, ja[0][-1], dtype=np.int32)]), ja[1], ja[2])
    jb2 = (np.concatenate([np.zeros(n, dtype=np.int32), jl[0]]), jl[1], 0.1 * jl[2])
    return A, B, ja2, jb2


def test_generalized_gradient_dense_and_fd():
    """d(sum of the lowest 16 eigenvalues of A x = lambda B x)/d theta from
    GCGE_HIP_EigGradientGeneralized, x_i^T (dA/dtheta - lambda_i dB/dtheta)
    x_i with B-normalized eigenvectors from the device solver, against the
    same formula with eigenvectors from dense LAPACK (scipy.linalg.eigh(A,
    B)) and against FP64 central differences of the dense loss, for
    parameters of A and of B; 1e-6 relative; non-degenerate spectrum
    enforced."""
    from gcge_hip import eig_gradient_generalized
    m = 32
    n = m * m
    kappa = P.vardiff2d_theta(m, var=0.25, seed=31)
    rho = P.vardiff2d_theta(m, var=0.25, seed=32)
    A, B, ja, jb = laplace_plus_mass(m, kappa, rho)
    w, V = sla.eigh(A.toarray(), B.toarray(), subset_by_index=[0, NEV])
    assert_simple(w)
    r = solve(A, B, NEV, X0=P.random_x0(n, NEV, 33), conv_type="NS", ev_tol=1e-11, ev_max_it=2000)
    assert r["num_unlock"] == 0
    np.testing.assert_allclose(r["eval"], w[:NEV], rtol=1e-10)
    X = r["evec"]
    np.testing.assert_allclose(X.T @ (B @ X), np.eye(NEV), atol=1e-10)
    mA, mB = Matrix(A), Matrix(B)
    g_dev = eig_gradient_generalized(mA, mB, X, r["eval"], np.ones(NEV), 2 * n, ja, jb)
    mA.close()
    mB.close()

    def sampled(M, Vv, wts):
        rows = np.repeat(np.arange(n), np.diff(M.indptr))
        ge = np.zeros(M.nnz)
        for i in range(NEV):
            ge += wts[i] * Vv[rows, i] * Vv[M.indices, i]
        return ge

    def gather(j, ge):
        jrp, jnz, jval = j
        return np.array([np.dot(jval[jrp[c]:jrp[c + 1]], ge[jnz[jrp[c]:jrp[c + 1]]]) for c in range(2 * n)])

    g_dense = gather(ja, sampled(A, V, np.ones(NEV))) - gather(jb, sampled(B, V, w[:NEV]))
    scale = np.abs(g_dense).max()
    np.testing.assert_allclose(g_dev, g_dense, rtol=1e-6, atol=1e-9 * scale)
    assert np.abs(g_dense[n:]).max() > 1e-3 * scale  # the B term is exercised

    def loss(k2, r2):
        A2, B2, _, _ = laplace_plus_mass(m, k2, r2)
        return float(np.sum(sla.eigh(A2.toarray(), B2.toarray(), eigvals_only=True, subset_by_index=[0, NEV - 1])))

    rng = np.random.default_rng(34)
    h = 1e-6
    for _ in range(3):
        v = rng.uniform(0.0, 1.0, 2 * n)
        vk, vr = v[:n].reshape(m, m), v[n:].reshape(m, m)
        fd = (loss(kappa + h * vk, rho + h * vr) - loss(kappa - h * vk, rho - h * vr)) / (2 * h)
        ad = float(np.dot(g_dev, v))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    # B's parameters alone (the -lambda dB/dtheta term)
    v = rng.uniform(0.0, 1.0, n).reshape(m, m)
    fd = (loss(kappa, rho + h * v) - loss(kappa, rho - h * v)) / (2 * h)
    ad = float(np.dot(g_dev[n:], v.ravel()))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
