"""Krylov mode: restarted GMRES preconditioned by the direct solver, for a
second-order Helmholtz operator whose wave number varies along x as well as
y (2D), or along x, y and z (3D), which the DST cannot diagonalise.

Oracle: a dense numpy solve of the same operator, assembled from the
second-order stencil vectors (tests/dense_ops.py) plus the per-point
diagonal h^2 (k^2(x, y) - mean_x k^2).  Bars: relative max-norm error
<= 1e-9 (GMRES tolerance 1e-12 times the operator's condition number), and
the preconditioned iteration counts stay small (21-93 measured) while
unpreconditioned GMRES has not converged after three times as many.
"""
import sys

import numpy as np
import pytest

from dense_ops import assemble2d

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def H():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    from hackathon_amd import helmholtz
    return helmholtz


# (grid, base wave number, relative variation of k along x, iteration bar)
@pytest.mark.parametrize("N,k0,amp,maxit", [(31, 12.0, 0.2, 40), (63, 20.0, 0.3, 150), (63, 8.0 + 2.0j, 0.25, 150)])
def test_gmres_variable_k(H, N, k0, amp, maxit):
    import torch
    from hackathon_amd import krylov
    say(f"GMRES variable k N={N} k0={k0} amp={amp}")
    ctx = H.Context(0)
    h = 1.0 / (N + 1)
    x = np.arange(1, N + 1) * h
    X, Y = np.meshgrid(x, x)                    # [j][i]
    k = k0 * (1.0 + amp * np.sin(np.pi * X) * (1.0 + Y))
    k2 = (k * k).astype(np.complex128)
    op = krylov.VariableK2D(ctx, N, k2)
    # dense operator: the preconditioner's stencil plus the per-point diagonal
    kbar = k2.mean(axis=1)
    coef = H.coefficients2d(2, N, N, h, h, np.concatenate([[kbar[0]], kbar, [kbar[-1]]]))
    A = assemble2d(N, N, coef) + np.diag((h * h * (k2 - kbar[:, None])).ravel())
    rng = np.random.default_rng(N)
    F = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    ref = np.linalg.solve(A, F)
    Ft = torch.from_numpy(F.reshape(N, N)).cuda()
    # the operator apply agrees with the dense operator
    u = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    Au = op.apply(torch.from_numpy(u.reshape(N, N)).cuda()).cpu().numpy().ravel()
    assert np.abs(Au - A @ u).max() <= 1e-12 * np.abs(A @ u).max()
    sol, info = op.solve(Ft, tol=1e-12)
    e = np.abs(sol.cpu().numpy().ravel() - ref).max() / np.abs(ref).max()
    say(f"  {info}, rel err vs dense solve {e:.2e}")
    assert info["converged"] and info["iterations"] <= maxit
    assert e <= 1e-9
    # without the preconditioner the same tolerance takes far longer
    _, plain = krylov.gmres(op.apply, Ft, None, tol=1e-12, restart=30, maxiter=3 * info["iterations"])
    say(f"  unpreconditioned after {plain['iterations']} iterations: rel residual {plain['rel_residual']:.2e}")
    assert not plain["converged"]


# (grid, base wave number, relative variation of k in x, y and z, iteration bar)
@pytest.mark.parametrize("N,k0,amp,maxit", [(15, 9.0, 0.3, 60), (15, 6.0 + 1.5j, 0.25, 60)])
def test_gmres_variable_k_3d(H, N, k0, amp, maxit):
    """3D, k(x, y, z): GMRES preconditioned by the 3D direct solve for k^2
    averaged over each z layer, against a dense numpy solve."""
    import torch
    from dense_ops import assemble3d
    from hackathon_amd import krylov
    say(f"GMRES 3D variable k N={N} k0={k0} amp={amp}")
    ctx = H.Context(0)
    h = 1.0 / (N + 1)
    x = np.arange(1, N + 1) * h
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")       # [l][j][i]
    k = k0 * (1.0 + amp * np.sin(np.pi * X) * np.cos(np.pi * Y) * (1.0 + Z))
    k2 = (k * k).astype(np.complex128)
    op = krylov.VariableK3D(ctx, N, k2)
    kbar = k2.mean(axis=(1, 2))
    coef = H.coefficients3d(2, N, N, N, h, h, h, np.concatenate([[kbar[0]], kbar, [kbar[-1]]]))
    A = assemble3d(N, N, N, np.concatenate(coef)) + np.diag((h * h * (k2 - kbar[:, None, None])).ravel())
    rng = np.random.default_rng(N + 3)
    F = rng.standard_normal(N ** 3) + 1j * rng.standard_normal(N ** 3)
    ref = np.linalg.solve(A, F)
    u = rng.standard_normal(N ** 3) + 1j * rng.standard_normal(N ** 3)
    Au = op.apply(torch.from_numpy(u.reshape(N, N, N)).cuda()).cpu().numpy().ravel()
    assert np.abs(Au - A @ u).max() <= 1e-12 * np.abs(A @ u).max()
    Ft = torch.from_numpy(F.reshape(N, N, N)).cuda()
    sol, info = op.solve(Ft, tol=1e-12)
    e = np.abs(sol.cpu().numpy().ravel() - ref).max() / np.abs(ref).max()
    say(f"  {info}, rel err vs dense solve {e:.2e}")
    assert info["converged"] and info["iterations"] <= maxit
    assert e <= 1e-9
    _, plain = krylov.gmres(op.apply, Ft, None, tol=1e-12, restart=30, maxiter=3 * info["iterations"])
    say(f"  unpreconditioned after {plain['iterations']} iterations: rel residual {plain['rel_residual']:.2e}")
    assert not plain["converged"]
