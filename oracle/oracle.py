"""ctypes wrapper of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke check and
the CPU-baseline leg of the benchmark, never by the product package.
The C sources restate the reference's 2D_Program and 3D_Program host
pipelines; see oracle.h for the per-function citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

SECOND, FOURTH, SIXTH = 2, 4, 6
REF, SINE = 0, 1


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        d = ctypes.c_double
        _lib.or_normInf.restype = d
        _lib.or_normL2.restype = d
        _lib.or_wtime.restype = d
        _lib.or_dst_plan.restype = ctypes.c_void_p
        _lib.or_dst_plan_free.argtypes = [ctypes.c_void_p]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _params(p):
    arr = np.zeros(8, dtype=np.float64)
    if p is not None:
        arr[: len(p)] = p
    return arr


def threads_default():
    return max(1, os.cpu_count() or 1)


def dst1_rows(x, threads=1):
    """Orthonormal DST-I along the last axis of a complex array."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    N = x.shape[-1]
    out = np.empty_like(x)
    lib().or_dst1_rows(ctypes.c_int(N), ctypes.c_long(x.size // N), _p(x), _p(out), ctypes.c_int(threads))
    return out


# ------------------------------------------------------------------ 2D
def domain2d(kind, p=None):
    dom = np.zeros(4)
    lib().or2d_domain(kind, _p(_params(p)), _p(dom))
    return dom


def setup2d(order, Nx, Ny, kind, p=None, threads=1):
    """coefficients + rhs: returns dict(coef=[bp,ap,b,a,bm,am], rhs, u)."""
    co = [np.zeros(Ny, np.complex128) for _ in range(6)]
    rhs = np.zeros(Nx * Ny, np.complex128)
    u = np.zeros(Nx * Ny, np.complex128)
    rc = lib().or2d_setup(order, Nx, Ny, kind, _p(_params(p)), *[_p(c) for c in co], _p(rhs), _p(u), threads)
    if rc != 0:
        raise ValueError(f"or2d_setup failed ({rc})")
    return {"coef": co, "rhs": rhs, "u": u}


def lu2d(Nx, Ny, coef, threads=1):
    L = np.zeros(Nx * Ny, np.complex128)
    U = np.zeros_like(L)
    Up = np.zeros_like(L)
    lib().or2d_LU(Nx, Ny, *[_p(c) for c in coef], _p(L), _p(U), _p(Up), threads)
    return L, U, Up


def solver2d(Nx, Ny, rhs, LUU, threads=1):
    sol = np.zeros(Nx * Ny, np.complex128)
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    lib().or2d_solver(Nx, Ny, _p(rhs), *[_p(a) for a in LUU], _p(sol), threads)
    return sol


class DstPlan:
    """Caller-held DST-I plan of length N (steady-state timing: plan creation outside the solve)."""

    def __init__(self, N):
        self.N = N
        self.h = ctypes.c_void_p(lib().or_dst_plan(N))

    def __del__(self):
        try:
            lib().or_dst_plan_free(self.h)
        except Exception:
            pass


def solver2d_planned(plan, Nx, Ny, rhs, LUU, threads=1):
    sol = np.zeros(Nx * Ny, np.complex128)
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    lib().or2d_solver_planned(plan.h, Nx, Ny, _p(rhs), *[_p(a) for a in LUU], _p(sol), threads)
    return sol


def residual2d(Nx, Ny, coef, sol, rhs):
    res = np.zeros(Nx * Ny, np.complex128)
    lib().or2d_residual(Nx, Ny, *[_p(c) for c in coef], _p(np.ascontiguousarray(sol)),
                        _p(np.ascontiguousarray(rhs)), _p(res))
    return res


def fields2d(Nx, Ny, kind, p=None):
    X, Y = Nx + 2, Ny + 2
    f_e = np.zeros(X * Y, np.complex128)
    u_e = np.zeros_like(f_e)
    fy_b = np.zeros(2 * Nx, np.complex128)
    fxx_b = np.zeros(2 * X + 2 * Y, np.complex128)
    fyy_b = np.zeros_like(fxx_b)
    k2_e = np.zeros(Y, np.complex128)
    k2y_b = np.zeros(2, np.complex128)
    lib().or2d_fields(Nx, Ny, kind, _p(_params(p)), _p(f_e), _p(u_e), _p(fy_b), _p(fxx_b), _p(fyy_b),
                      _p(k2_e), _p(k2y_b))
    return dict(f_e=f_e, u_e=u_e, fy_b=fy_b, fxx_b=fxx_b, fyy_b=fyy_b, k2_e=k2_e, k2y_b=k2y_b)


def solve2d(order, N, kind, p=None, threads=1, Ny=None):
    """Full reference pipeline: setup, LU, solver, residual."""
    Ny = N if Ny is None else Ny
    s = setup2d(order, N, Ny, kind, p, threads)
    LUU = lu2d(N, Ny, s["coef"], threads)
    sol = solver2d(N, Ny, s["rhs"], LUU, threads)
    res = residual2d(N, Ny, s["coef"], sol, s["rhs"])
    s.update(sol=sol, res=res, LU=LUU)
    return s


# ------------------------------------------------------------------ 3D
def domain3d(kind, p=None):
    dom = np.zeros(6)
    lib().or3d_domain(kind, _p(_params(p)), _p(dom))
    return dom


def setup3d(order, Nx, Ny, Nz, kind, p=None, threads=1):
    coef = np.zeros(12 * Nz, np.complex128)
    n = Nx * Ny * Nz
    rhs = np.zeros(n, np.complex128)
    u = np.zeros(n, np.complex128)
    rc = lib().or3d_setup(order, Nx, Ny, Nz, kind, _p(_params(p)), _p(coef), _p(rhs), _p(u), threads)
    if rc != 0:
        raise ValueError(f"or3d_setup failed ({rc})")
    return {"coef": coef, "rhs": rhs, "u": u}


def lu3d(Nx, Ny, Nz, coef):
    n = Nx * Ny * Nz
    L = np.zeros(n, np.complex128)
    U = np.zeros_like(L)
    Up = np.zeros_like(L)
    lib().or3d_LU(Nx, Ny, Nz, _p(coef), _p(L), _p(U), _p(Up))
    return L, U, Up


def solver3d(Nx, Ny, Nz, rhs, LUU, threads=1):
    sol = np.zeros(Nx * Ny * Nz, np.complex128)
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    lib().or3d_solver(Nx, Ny, Nz, _p(rhs), *[_p(a) for a in LUU], _p(sol), threads)
    return sol


def solver3d_planned(px, py, Nx, Ny, Nz, rhs, LUU, threads=1):
    sol = np.zeros(Nx * Ny * Nz, np.complex128)
    rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
    lib().or3d_solver_planned(px.h, py.h, Nx, Ny, Nz, _p(rhs), *[_p(a) for a in LUU], _p(sol), threads)
    return sol


def residual3d(Nx, Ny, Nz, coef, sol, rhs):
    res = np.zeros(Nx * Ny * Nz, np.complex128)
    lib().or3d_residual(Nx, Ny, Nz, _p(coef), _p(np.ascontiguousarray(sol)), _p(np.ascontiguousarray(rhs)),
                        _p(res))
    return res


def fields3d(Nx, Ny, Nz, kind, p=None):
    X, Y, Z = Nx + 2, Ny + 2, Nz + 2
    f_e = np.zeros(X * Y * Z, np.complex128)
    u_e = np.zeros_like(f_e)
    fxx_f = np.zeros(2 * Y * Z, np.complex128)
    fyy_f = np.zeros(2 * X * Z, np.complex128)
    fzz_f = np.zeros(2 * X * Y, np.complex128)
    k2_e = np.zeros(Z, np.complex128)
    k2z_b = np.zeros(4, np.complex128)
    lib().or3d_fields(Nx, Ny, Nz, kind, _p(_params(p)), _p(f_e), _p(u_e), _p(fxx_f), _p(fyy_f), _p(fzz_f),
                      _p(k2_e), _p(k2z_b))
    return dict(f_e=f_e, u_e=u_e, fxx_f=fxx_f, fyy_f=fyy_f, fzz_f=fzz_f, k2_e=k2_e, k2z_b=k2z_b)


def solve3d(order, N, kind, p=None, threads=1, Ny=None, Nz=None):
    Ny = N if Ny is None else Ny
    Nz = N if Nz is None else Nz
    s = setup3d(order, N, Ny, Nz, kind, p, threads)
    LUU = lu3d(N, Ny, Nz, s["coef"])
    sol = solver3d(N, Ny, Nz, s["rhs"], LUU, threads)
    res = residual3d(N, Ny, Nz, s["coef"], sol, s["rhs"])
    s.update(sol=sol, res=res, LU=LUU)
    return s


def normInf(x):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    return lib().or_normInf(ctypes.c_long(x.size), _p(x))


def normL2(x):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    return lib().or_normL2(ctypes.c_long(x.size), _p(x))


def wtime():
    return lib().or_wtime()
