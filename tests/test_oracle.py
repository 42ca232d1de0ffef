"""Pins of the CPU FP64 oracle (oracle/), each against something that does
not come from the oracle itself:

  or_dst1_rows        scipy golden vectors (tests/golden/dst1_golden.npz),
                      O(N^2) brute-force sine sums, S*S = I and the
                      eigen-decomposition of tridiag(1,0,1)
  or2d_LU/or2d_solver dense numpy solve of the assembled 9-point operator
  or3d_LU/or3d_solver dense numpy solve of the assembled 27-point operator
  or2d_setup (rhs)    2nd/4th/6th-order convergence against the analytic
                      solution (ratios 4/16/64 per halving of h)
  or3d_setup (rhs)    paper Tables 4-6 known answers (tests/golden/paper_tables.json)
                      and the convergence ratios
  or*_residual        the residual of a dense solve is ~1e-14 while a
                      planted stencil error is not
"""
import json
import os

import numpy as np
import pytest

from dense_ops import assemble2d, assemble3d

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def dst_matrix(n):
    k = np.arange(1, n + 1)
    return np.sqrt(2.0 / (n + 1)) * np.sin(np.pi * np.outer(k, k) / (n + 1))


def test_dst_golden(oracle):
    g = np.load(os.path.join(GOLD, "dst1_golden.npz"))
    for key in g.files:
        if not key.startswith("x_"):
            continue
        n = key[2:]
        y = oracle.dst1_rows(g[key])
        np.testing.assert_allclose(y, g["y_" + n], rtol=0, atol=1e-13 * max(1.0, np.abs(g["y_" + n]).max()))


@pytest.mark.parametrize("n", [1, 4, 9, 100, 126, 201])
def test_dst_bruteforce_and_invariants(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    S = dst_matrix(n)
    y = oracle.dst1_rows(x)
    np.testing.assert_allclose(y, x @ S.T, atol=1e-12)
    # self-inverse
    np.testing.assert_allclose(oracle.dst1_rows(y), x, atol=1e-12)
    # S diagonalises tridiag(1,0,1) with eigenvalues 2cos(k pi/(n+1))
    T = np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    lam = 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    np.testing.assert_allclose(S @ T @ S, np.diag(lam), atol=1e-12)


@pytest.mark.parametrize("order,kind,p,Nx,Ny", [
    (2, 0, None, 7, 5), (4, 0, None, 6, 9), (6, 0, None, 8, 8),
    (4, 1, [10.0, 0.5, 3, 2, 0.7], 9, 7), (6, 1, [50.0, 0.0, 3, 2, 1.1], 11, 11),
])
def test_solver2d_dense(oracle, order, kind, p, Nx, Ny):
    s = oracle.setup2d(order, Nx, Ny, kind, p)
    LUU = oracle.lu2d(Nx, Ny, s["coef"])
    sol = oracle.solver2d(Nx, Ny, s["rhs"], LUU)
    A = assemble2d(Nx, Ny, s["coef"])
    ref = np.linalg.solve(A, s["rhs"])
    assert np.abs(sol - ref).max() <= 1e-12 * np.abs(ref).max()
    res = oracle.residual2d(Nx, Ny, s["coef"], sol, s["rhs"])
    np.testing.assert_allclose(res, A @ sol - s["rhs"], atol=1e-13 * np.abs(s["rhs"]).max())


@pytest.mark.parametrize("order,kind,p,N", [
    (2, 0, [10, 9, 10, 9], (4, 5, 6)), (4, 0, [10, 9, 10, 9], (5, 4, 5)), (6, 0, [10, 9, 10, 9], (5, 5, 5)),
    (6, 1, [30.0, 0.0, 3, 2, 1, 1.0, 1.0], (6, 5, 4)),
])
def test_solver3d_dense(oracle, order, kind, p, N):
    Nx, Ny, Nz = N
    s = oracle.setup3d(order, Nx, Ny, Nz, kind, p)
    LUU = oracle.lu3d(Nx, Ny, Nz, s["coef"])
    sol = oracle.solver3d(Nx, Ny, Nz, s["rhs"], LUU)
    A = assemble3d(Nx, Ny, Nz, s["coef"])
    ref = np.linalg.solve(A, s["rhs"])
    assert np.abs(sol - ref).max() <= 1e-12 * np.abs(ref).max()
    res = oracle.residual3d(Nx, Ny, Nz, s["coef"], sol, s["rhs"])
    np.testing.assert_allclose(res, A @ sol - s["rhs"], atol=1e-13 * np.abs(s["rhs"]).max())


@pytest.mark.parametrize("order,ratio", [(2, 4.0), (4, 16.0), (6, 64.0)])
@pytest.mark.parametrize("kind,p", [(0, None), (1, [10.0, 0.0, 3, 2, 0.3])])
def test_convergence2d(oracle, order, ratio, kind, p):
    errs = []
    for N in (31, 63, 127):
        s = oracle.solve2d(order, N, kind, p, threads=4)
        errs.append(oracle.normInf(s["sol"] - s["u"]))
        assert oracle.normInf(s["res"]) < 1e-12
    r = errs[-2] / errs[-1]
    assert 0.85 * ratio < r < 1.2 * ratio, errs


@pytest.mark.parametrize("order,ratio", [(4, 16.0), (6, 64.0)])
def test_convergence3d_sine(oracle, order, ratio):
    p = [20.0, 0.0, 3, 2, 1, 1.0, 1.0]
    errs = []
    for N in (15, 31, 63):
        s = oracle.solve3d(order, N, 1, p, threads=8)
        errs.append(oracle.normInf(s["sol"] - s["u"]))
    r = errs[-2] / errs[-1]
    assert 0.85 * ratio < r < 1.25 * ratio, errs


@pytest.mark.parametrize("order", [2, 4, 6])
def test_paper_known_answers_125(oracle, order):
    t = json.load(open(os.path.join(GOLD, "paper_tables.json")))
    s = oracle.solve3d(order, 125, 0, t["params"], threads=8)
    err = oracle.normInf(s["sol"] - s["u"])
    want = t["max_err"]["125"][str(order)]
    # the paper prints 8 significant digits
    assert abs(err - want) <= 1e-7 * want, (err, want)
    assert oracle.normL2(s["res"]) < 1e-11


def test_planted_stencil_error_is_caught(oracle):
    """A wrong coefficient (a dropped term) in the operator must show up in
    the dense check that pins the solver: perturb one coefficient vector."""
    Nx = Ny = 8
    s = oracle.setup2d(6, Nx, Ny, 1, [30.0, 0.0, 2, 3, 0.0])
    coef = [c.copy() for c in s["coef"]]
    A = assemble2d(Nx, Ny, coef)
    ref = np.linalg.solve(A, s["rhs"])
    bad = [c.copy() for c in coef]
    bad[0] = bad[0] * 0  # drop the corner weights bp
    sol = oracle.solver2d(Nx, Ny, s["rhs"], oracle.lu2d(Nx, Ny, bad))
    assert np.abs(sol - ref).max() > 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("N", [63, 100])
def test_planned_solvers_equal_unplanned(oracle, N):
    """The planned solvers (the DST plan created once by the caller, as in the
    steady-state CPU baseline) give the same bits as the solvers that plan
    inside the call, over several reuses of one plan; 2D and 3D, FFT and
    Bluestein lengths."""
    s = oracle.solve2d(6, N, 1, [10.0, 0.0, 3, 2, 0.3], threads=2)
    plan = oracle.DstPlan(N)
    for _ in range(2):
        assert np.array_equal(oracle.solver2d_planned(plan, N, N, s["rhs"], s["LU"], 2), s["sol"])
    n3 = 15 if N == 63 else 20
    t = oracle.setup3d(6, n3, n3, n3, oracle.SINE, [10.0, 0.0, 1, 2, 3, 1.0, 1.0], threads=2)
    LUU = oracle.lu3d(n3, n3, n3, t["coef"])
    ref = oracle.solver3d(n3, n3, n3, t["rhs"], LUU, 2)
    p3 = oracle.DstPlan(n3)
    for _ in range(2):
        assert np.array_equal(oracle.solver3d_planned(p3, p3, n3, n3, n3, t["rhs"], LUU, 2), ref)
