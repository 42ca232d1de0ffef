"""Robin / Sommerfeld closure of the tridiagonal axis (hh_robin2d /
hh_robin3d, host functions of the C ABI), on the CPU.

No reference counterpart: the reference solves Dirichlet problems only
(2D_Program/main_functions/rhs.c:126-148, 3D_Program/main_functions/
rhs.c:167-174).  Oracle: the dense system with explicit ghost rows and
boundary equations of robin_dense.py; the folded band vectors, assembled
by dense_ops.py, plus the right-hand-side corrections must give the same
interior solution.  A fold that drops the 2h factor or a sign, moves a band
to the wrong side, or misses the neighbour sums of the correction fails.
Bar: 1e-12 relative in the max norm.
"""
import numpy as np
import pytest

from dense_ops import assemble2d, assemble3d
from robin_dense import folded_rhs, solve_ghost2d, solve_ghost3d


@pytest.fixture(scope="module")
def H():
    from hackathon_amd import helmholtz
    return helmholtz


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


SIDES = [("lo", True, False), ("hi", False, True), ("both", True, True)]


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("name,use_lo,use_hi", SIDES)
def test_robin2d_matches_ghost_system(H, order, name, use_lo, use_hi):
    rng = np.random.default_rng(order * 7 + len(name))
    Nx, Ny = 5, 6
    h = 1.0 / (Nx + 1)
    k = 9.0
    coef = H.coefficients2d(order, Nx, Ny, h, h, np.full(Ny + 2, k * k + 0j))
    lo = (complex(*rng.standard_normal(2)), crand(rng, Nx)) if use_lo else None
    hi = (-1j * k, crand(rng, Nx)) if use_hi else None
    folded, corr = H.robin2d(coef, Nx, Ny, h, lo, hi)
    F = crand(rng, Ny, Nx)
    ref = solve_ghost2d(coef, Nx, Ny, h, F, lo, hi)
    A = assemble2d(Nx, Ny, folded)
    u = np.linalg.solve(A, folded_rhs(F, corr, Ny).ravel()).reshape(Ny, Nx)
    assert relerr(u, ref) <= 1e-12
    # the caller's vectors are left untouched
    again, _ = H.robin2d(coef, Nx, Ny, h, lo, hi)
    for x, y in zip(folded, again):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("name,use_lo,use_hi", SIDES)
def test_robin3d_matches_ghost_system(H, order, name, use_lo, use_hi):
    rng = np.random.default_rng(order * 11 + len(name))
    Nx, Ny, Nz = 3, 4, 5
    h = 1.0 / (Nx + 1)
    k = 6.0
    coef = H.coefficients3d(order, Nx, Ny, Nz, 1.0 / (Nx + 1), 1.0 / (Ny + 1), h, np.full(Nz + 2, k * k + 0j))
    lo = (-1j * k, crand(rng, Ny, Nx)) if use_lo else None
    hi = (complex(*rng.standard_normal(2)), crand(rng, Ny, Nx)) if use_hi else None
    folded, corr = H.robin3d(coef, Nx, Ny, Nz, h, lo, hi)
    F = crand(rng, Nz, Ny, Nx)
    ref = solve_ghost3d(coef, Nx, Ny, Nz, h, F, lo, hi)
    A = assemble3d(Nx, Ny, Nz, np.stack(folded))
    u = np.linalg.solve(A, folded_rhs(F, corr, Nz).ravel()).reshape(Nz, Ny, Nx)
    assert relerr(u, ref) <= 1e-12


def test_robin_zero_data_and_dirichlet_sides(H):
    """g = None gives zero corrections; a side given as None is left as it was."""
    Nx, Ny = 4, 5
    h = 0.2
    coef = H.coefficients2d(4, Nx, Ny, h, h, np.full(Ny + 2, 16.0 + 0j))
    folded, corr = H.robin2d(coef, Nx, Ny, h, lo=(-4j, None))
    assert corr[1] is None and np.array_equal(corr[0], np.zeros(Nx, np.complex128))
    for v, w in zip(folded, coef):
        assert np.array_equal(v[1:], w[1:])
    assert folded[4][0] == 0 and folded[5][0] == 0          # minus bands of the lo row vanish


def test_robin_errors(H):
    coef = H.coefficients2d(2, 4, 1, 0.2, 0.2, np.full(3, 1.0 + 0j))
    with pytest.raises(H.HHError):
        H.robin2d(coef, 4, 1, 0.2, lo=(1.0, None))             # one row cannot hold a Robin side
    coef = H.coefficients2d(2, 4, 4, 0.2, 0.2, np.full(6, 1.0 + 0j))
    with pytest.raises(H.HHError):
        H.robin2d(coef, 4, 4, 0.0, lo=(1.0, None))             # h must be positive
