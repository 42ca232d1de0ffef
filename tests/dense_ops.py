"""Dense assembly of the compact operators from their coefficient vectors
(the 9-point stencil of 2D_Program/main_functions/residual.c:19-29 and the
27-point stencil of 3D_Program/main_functions/residual.c:69-82), for
direct numpy solves on small grids."""
import numpy as np


def assemble2d(Nx, Ny, coef):
    bp, ap, b, a, bm, am = coef
    n = Nx * Ny
    A = np.zeros((n, n), np.complex128)
    for j in range(Ny):
        for i in range(Nx):
            r = i + j * Nx
            w = {(-1, 1): bp[j], (0, 1): ap[j], (1, 1): bp[j], (-1, 0): b[j], (0, 0): a[j], (1, 0): b[j],
                 (-1, -1): bm[j], (0, -1): am[j], (1, -1): bm[j]}
            for (di, dj), v in w.items():
                ii, jj = i + di, j + dj
                if 0 <= ii < Nx and 0 <= jj < Ny:
                    A[r, ii + jj * Nx] += v
    return A


def assemble3d(Nx, Ny, Nz, coef):
    c = coef.reshape(12, Nz)
    a, b, cc, d, ap, bp, cp, dp, am, bm, cm, dm = c
    n = Nx * Ny * Nz
    A = np.zeros((n, n), np.complex128)
    for l in range(Nz):
        lev = {-1: (am, bm, cm, dm), 0: (a, b, cc, d), 1: (ap, bp, cp, dp)}
        for j in range(Ny):
            for i in range(Nx):
                r = i + j * Nx + l * Nx * Ny
                for dl, (A_, B_, C_, D_) in lev.items():
                    w = {(0, 0): D_[l], (1, 0): B_[l], (-1, 0): B_[l], (0, 1): C_[l], (0, -1): C_[l],
                         (1, 1): A_[l], (-1, 1): A_[l], (1, -1): A_[l], (-1, -1): A_[l]}
                    for (di, dj), v in w.items():
                        ii, jj, ll = i + di, j + dj, l + dl
                        if 0 <= ii < Nx and 0 <= jj < Ny and 0 <= ll < Nz:
                            A[r, ii + jj * Nx + ll * Nx * Ny] += v
    return A
