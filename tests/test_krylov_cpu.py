"""CPU checks of the restarted GMRES in hackathon_amd/krylov.py (the
Krylov mode's outer iteration; the GPU test covers the operator and the
preconditioner): complex systems with a known solution from numpy."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def krylov():
    from hackathon_amd import krylov
    return krylov


def system(n, scale, seed):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 4 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * scale
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return A, b, np.linalg.solve(A, b)


@pytest.mark.parametrize("restart", [5, 20, 200])
def test_gmres_restarts_converge(krylov, restart):
    import torch
    A, b, ref = system(200, 0.1, 1)
    At = torch.from_numpy(A)
    x, info = krylov.gmres(lambda v: At @ v, torch.from_numpy(b), None, tol=1e-12, restart=restart, maxiter=2000)
    assert info["converged"] and info["rel_residual"] <= 1e-12
    assert np.abs(x.numpy() - ref).max() <= 1e-11 * np.abs(ref).max()


def test_gmres_preconditioned(krylov):
    """A near-inverse right preconditioner cuts the iterations to a few."""
    import torch
    A, b, ref = system(200, 0.2, 2)
    At, bt = torch.from_numpy(A), torch.from_numpy(b)
    Mt = torch.from_numpy(np.linalg.inv(A + 0.05 * np.eye(200)))
    x, info = krylov.gmres(lambda v: At @ v, bt, lambda v: Mt @ v, tol=1e-12, restart=10, maxiter=100)
    assert info["converged"] and info["iterations"] <= 12
    assert np.abs(x.numpy() - ref).max() <= 1e-11 * np.abs(ref).max()
    _, plain = krylov.gmres(lambda v: At @ v, bt, None, tol=1e-12, restart=10, maxiter=3 * info["iterations"])
    assert not plain["converged"]


def test_gmres_zero_rhs(krylov):
    import torch
    x, info = krylov.gmres(lambda v: v, torch.zeros(8, dtype=torch.complex128))
    assert info["converged"] and info["iterations"] == 0 and float(x.abs().max()) == 0.0
