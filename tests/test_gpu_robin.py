"""GPU solves with a Robin / Sommerfeld boundary on the tridiagonal axis:
the band vectors and right-hand side folded by hh_robin2d / hh_robin3d go
through the unchanged plans (DST, mode solves, eigenvalue assembly).

No reference counterpart (the reference solves Dirichlet problems only).
Oracles: (1) the dense system with explicit ghost rows and boundary
equations (robin_dense.py), bar 1e-12 relative; (2) a manufactured
solution of the second-order scheme with a Sommerfeld-type condition at
y = 0, whose error must fall as h^2 (observed order >= 1.9 per halving).
"""
import sys

import numpy as np
import pytest

from robin_dense import folded_rhs, solve_ghost2d, solve_ghost3d

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


@pytest.fixture(scope="module")
def ctx(H):
    return H.Context(0)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# (order, Nx, Ny, k, lo alpha, hi alpha); None = Dirichlet side; "s" = Sommerfeld -ik
CASES2D = [(4, 15, 16, 20.0, "s", None), (6, 31, 20, 12.0, 0.0, 0.5), (2, 15, 9, 8.0, None, "s"),
           (6, 63, 12, 30.0, "s", "s")]


@pytest.mark.parametrize("order,Nx,Ny,k,alo,ahi", CASES2D)
def test_robin2d_matches_ghost_system(H, ctx, order, Nx, Ny, k, alo, ahi):
    rng = np.random.default_rng(Nx * 100 + Ny)
    h = 1.0 / (Nx + 1)
    coef = H.coefficients2d(order, Nx, Ny, h, h, np.full(Ny + 2, k * k + 0j))
    side = lambda a: None if a is None else ((-1j * k) if a == "s" else a, crand(rng, Nx))
    lo, hi = side(alo), side(ahi)
    folded, corr = H.robin2d(coef, Nx, Ny, h, lo, hi)
    F = [crand(rng, Ny, Nx) for _ in range(2)]
    plan = H.Plan2D(ctx, Nx, Ny, folded)
    say(f"2D Robin order={order} Nx={Nx} Ny={Ny} k={k} lo={alo} hi={ahi} real={plan.is_real_operator}")
    if alo != "s" and ahi != "s":
        assert plan.is_real_operator      # real alpha and k keep the real band arithmetic
    sol = plan.solve(dev(np.stack([folded_rhs(f, corr, Ny) for f in F]))).cpu().numpy()
    for b in range(2):
        ref = solve_ghost2d(coef, Nx, Ny, h, F[b], lo, hi)
        err = relerr(sol[b].reshape(Ny, Nx), ref)
        say(f"  rhs {b}: rel err vs ghost system {err:.2e}")
        assert err <= 1e-12


def test_robin2d_mixed_precision(H, ctx):
    """FP32 bulk + FP64 refinement on a folded operator: the refinement's
    residual uses the same folded bands, so two fixed sweeps land within
    1e-10 of the ghost-row system."""
    rng = np.random.default_rng(77)
    Nx, Ny, k = 63, 40, 25.0
    h = 1.0 / (Nx + 1)
    coef = H.coefficients2d(6, Nx, Ny, h, h, np.full(Ny + 2, k * k + 0j))
    lo, hi = (-1j * k, crand(rng, Nx)), (0.4, crand(rng, Nx))
    folded, corr = H.robin2d(coef, Nx, Ny, h, lo, hi)
    F = crand(rng, Ny, Nx)
    plan = H.Plan2D(ctx, Nx, Ny, folded, precision=H.PREC_MIXED)
    plan.set_refine(0.0, 2)
    sol = plan.solve(dev(folded_rhs(F, corr, Ny))).cpu().numpy().reshape(Ny, Nx)
    err = relerr(sol, solve_ghost2d(coef, Nx, Ny, h, F, lo, hi))
    say(f"2D Robin mixed Nx={Nx} Ny={Ny}: rel err vs ghost system {err:.2e}")
    assert err <= 1e-10


CASES3D = [(6, 7, 7, 8, 10.0, "s", "s"), (4, 7, 5, 6, 6.0, 0.3, None), (2, 9, 6, 7, 4.0, None, "s")]


@pytest.mark.parametrize("order,Nx,Ny,Nz,k,alo,ahi", CASES3D)
def test_robin3d_matches_ghost_system(H, ctx, order, Nx, Ny, Nz, k, alo, ahi):
    rng = np.random.default_rng(Nx * 1000 + Ny * 10 + Nz)
    hx, hy = 1.0 / (Nx + 1), 1.0 / (Ny + 1)
    h = hx
    coef = H.coefficients3d(order, Nx, Ny, Nz, hx, hy, h, np.full(Nz + 2, k * k + 0j))
    side = lambda a: None if a is None else ((-1j * k) if a == "s" else a, crand(rng, Ny, Nx))
    lo, hi = side(alo), side(ahi)
    folded, corr = H.robin3d(coef, Nx, Ny, Nz, h, lo, hi)
    F = crand(rng, Nz, Ny, Nx)
    plan = H.Plan3D(ctx, Nx, Ny, Nz, folded)
    say(f"3D Robin order={order} {Nx}x{Ny}x{Nz} k={k} lo={alo} hi={ahi}")
    sol = plan.solve(dev(folded_rhs(F, corr, Nz))).cpu().numpy().reshape(Nz, Ny, Nx)
    ref = solve_ghost3d(coef, Nx, Ny, Nz, h, F, lo, hi)
    err = relerr(sol, ref)
    say(f"  rel err vs ghost system {err:.2e}")
    assert err <= 1e-12


def test_robin2d_second_order_convergence(H, ctx):
    """u = (1 + i/2) sin(2 pi x) sin(1.5 pi (1 - y)) on the unit square, with
    du/dn - i k u = g at y = 0 (rows y_j = j h, j = 0..Ny-1) and u = 0 on the
    other sides; second-order scheme, F = h^2 (Delta u + k^2 u)."""
    k, a, c = 10.0, 2.0, 1.5
    amp = 1.0 + 0.5j
    errs = []
    for Ny in (64, 128, 256):
        h = 1.0 / Ny
        Nx = Ny - 1
        x = (np.arange(Nx) + 1) * h
        y = np.arange(Ny) * h
        X, Y = np.meshgrid(x, y)
        u = amp * np.sin(a * np.pi * X) * np.sin(c * np.pi * (1 - Y))
        f = (k * k - (a * np.pi) ** 2 - (c * np.pi) ** 2) * u
        alpha = -1j * k
        dudy0 = amp * np.sin(a * np.pi * x) * (-c * np.pi) * np.cos(c * np.pi)
        g = -dudy0 + alpha * amp * np.sin(a * np.pi * x) * np.sin(c * np.pi)
        coef = H.coefficients2d(2, Nx, Ny, h, h, np.full(Ny + 2, k * k + 0j))
        folded, corr = H.robin2d(coef, Nx, Ny, h, lo=(alpha, g))
        plan = H.Plan2D(ctx, Nx, Ny, folded)
        sol = plan.solve(dev(folded_rhs(h * h * f, corr, Ny))).cpu().numpy().reshape(Ny, Nx)
        errs.append(relerr(sol, u))
        say(f"Ny={Ny}: rel err vs exact {errs[-1]:.3e}")
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    say(f"observed orders {orders}")
    assert errs[-1] < 1e-3
    assert min(orders) >= 1.9
