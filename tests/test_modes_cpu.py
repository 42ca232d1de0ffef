"""Host-side mode classification and reference-order factors (no GPU).

The library splits every operator's modes into well-separated ones (Varah's
bound on the condition number of the mode's tridiagonal system at most
kappa) and near-resonant ones, whose LU factors it computes on the host in
the reference's operation order (2D_Program/main_functions/LU.c:16-40,
3D_Program/main_functions/LU.c:34-72).  These tests pin
  * the factors: bit-identical to the oracle's LU() for real operators
    (to 1e-15 relative for complex ones, whose complex division may round
    differently),
  * the classification: the flagged set equals an independent numpy
    evaluation of Varah's bound on the bands the oracle's LU() implies,
  * the singularity verdict: an operator whose k^2 is a discrete eigenvalue
    is singular, one next to it is not.
"""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def H():
    from hackathon_amd import helmholtz
    helmholtz.lib()
    return helmholtz


def bands2d(coef, N, i):
    bp, ap, b, a, bm, am = coef
    lam = 2 * np.cos((i + 1) * np.pi / (N + 1))
    sub, dia, sup = bm * lam + am, b * lam + a, bp * lam + ap
    sub = sub.copy(); sup = sup.copy()
    sub[0] = 0; sup[-1] = 0
    return sub, dia, sup


def varah_flags(coef, N, kappa):
    out = []
    for i in range(N):
        sub, dia, sup = bands2d(coef, N, i)
        margin = np.abs(dia) - np.abs(sub) - np.abs(sup)
        rows = np.abs(dia) + np.abs(sub) + np.abs(sup)
        if not margin.min() > 0 or rows.max() > kappa * margin.min():
            out.append(i)
    return out


@pytest.mark.parametrize("order,N,k", [(6, 63, 10.0), (4, 127, 30.0), (6, 255, 50.0), (2, 64, 7.3)])
def test_factors_bit_identical_2d(H, order, N, k):
    s = O.setup2d(order, N, N, O.SINE, [k, 0.0, 1, 1, 0.0])
    L, U, Up = O.lu2d(N, N, s["coef"])
    for i in (0, 1, N // 3, N - 1):
        f, sing = H.operator_mode_factors(2, N, N, 1, s["coef"], i)
        assert not sing
        sl = slice(i * N, (i + 1) * N)
        assert np.array_equal(f[:, 0].real, L[sl].real) and not f[:, 0].imag.any()
        assert np.array_equal(f[:, 1].real, 1.0 / U[sl].real)
        assert np.array_equal(f[:-1, 2].real, Up[sl][:-1].real) and f[-1, 2] == 0


def test_factors_complex_2d(H):
    # the reference's own 2D problem has a complex k(y) = y^2 + i y
    N = 47
    s = O.setup2d(6, N, N, O.REF)
    L, U, Up = O.lu2d(N, N, s["coef"])
    for i in (0, 5, N - 1):
        f, _ = H.operator_mode_factors(2, N, N, 1, s["coef"], i)
        sl = slice(i * N, (i + 1) * N)
        assert np.abs(f[:, 0] - L[sl]).max() <= 1e-15 * np.abs(L[sl]).max()
        assert np.abs(f[:, 1] * U[sl] - 1).max() <= 1e-15


@pytest.mark.parametrize("order,N,k", [(6, 15, 12.0), (4, 16, 20.0), (2, 12, 9.0)])
def test_factors_bit_identical_3d(H, order, N, k):
    s = O.setup3d(order, N, N, N, O.SINE, [k, 0.0, 1, 1, 1, 1.0, 1.0])
    L, U, Up = O.lu3d(N, N, N, s["coef"])
    coef = s["coef"].reshape(12, N)
    for (i, j) in ((0, 0), (3, 1), (N - 1, N - 2)):
        f, sing = H.operator_mode_factors(3, N, N, N, coef, i + j * N)
        assert not sing
        base = i * N + j * N * N          # LU.c layout: l + i*Nz + j*Nx*Nz
        sl = slice(base, base + N)
        assert np.array_equal(f[:, 0].real, L[sl].real)
        assert np.array_equal(f[:, 1].real, U[sl].real)
        assert np.array_equal(f[:-1, 2].real, Up[sl][:-1].real)


@pytest.mark.parametrize("order,N,k,kappa", [(6, 255, 50.0, 4096.0), (4, 511, 14.1, 4096.0),
                                             (6, 1023, 5.3, 4096.0), (6, 127, 30.0, 100.0)])
def test_classification_matches_varah(H, order, N, k, kappa):
    s = O.setup2d(order, N, N, O.SINE, [k, 0.0, 1, 1, 0.0])
    modes, sing, kap = H.operator_modes(2, N, N, 1, s["coef"], kappa)
    assert modes == varah_flags(s["coef"], N, kappa)
    assert not sing and np.isfinite(kap)
    # every indefinite mode (|lambda| below k^2) is among them
    assert len(modes) >= int(k / np.pi) - 1


def test_classification_3d_uniform_and_variable(H):
    # constant k (uniform rows) and the reference's k(z) = A - B sin(Cz): the
    # one-row shortcut for uniform tables must agree with the full scan
    N = 23
    for kind, p in ((O.SINE, [14.0, 0.0, 1, 1, 1, 1.0, 1.0]), (O.REF, [10.0, 9.0, 10.0, 9.0])):
        s = O.setup3d(6, N, N, N, kind, p)
        coef = s["coef"].reshape(12, N)
        modes, sing, _ = H.operator_modes(3, N, N, N, coef, 4096.0)
        lx = 2 * np.cos(np.arange(1, N + 1) * np.pi / (N + 1))
        ref = []
        for m in range(N * N):
            a, b = lx[m % N], lx[m // N]
            c = coef
            dia = c[0] * a * b + c[1] * a + c[2] * b + c[3]
            sup = c[4] * a * b + c[5] * a + c[6] * b + c[7]
            sub = c[8] * a * b + c[9] * a + c[10] * b + c[11]
            sub = sub.copy(); sup = sup.copy(); sub[0] = 0; sup[-1] = 0
            mg = np.abs(dia) - np.abs(sub) - np.abs(sup)
            rs = np.abs(dia) + np.abs(sub) + np.abs(sup)
            if not mg.min() > 0 or rs.max() > 4096.0 * mg.min():
                ref.append(m)
        assert modes == ref, kind
        assert not sing


def test_singular_at_eigenvalue(H):
    # second-order 2D scheme, constant k: the mode (i, j) eigenvalue of the
    # discrete operator is (lam_i - 2) + (2 cos((j+1) pi/(N+1)) - 2) + k^2 h^2;
    # k^2 chosen to zero it is singular, a slightly different k is not
    N = 31
    h = 1.0 / (N + 1)
    i, j = 2, 3
    k2 = (4 - 2 * np.cos((i + 1) * np.pi / (N + 1)) - 2 * np.cos((j + 1) * np.pi / (N + 1))) / h ** 2
    k2e = np.full(N + 2, k2, np.complex128)
    coef = H.coefficients2d(2, N, N, h, h, k2e)
    modes, sing, kap = H.operator_modes(2, N, N, 1, coef)
    assert sing and i in modes, (modes, kap)
    coef2 = H.coefficients2d(2, N, N, h, h, np.full(N + 2, k2 * (1 + 1e-3), np.complex128))
    modes2, sing2, kap2 = H.operator_modes(2, N, N, 1, coef2)
    assert not sing2 and kap2 < 1e14


def test_kappa_zero_flags_every_mode(H):
    N = 40
    s = O.setup2d(4, N, N, O.SINE, [3.0, 0.0, 1, 1, 0.0])
    modes, _, _ = H.operator_modes(2, N, N, 1, s["coef"], 0.0)
    assert modes == list(range(N))
    modes, _, _ = H.operator_modes(2, N, N, 1, s["coef"], 1e300)
    assert modes == []   # k = 3 < pi: every mode diagonally dominant
