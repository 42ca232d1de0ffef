"""Dense solves of compact-stencil problems with a Robin boundary
du/dn + alpha u = g on the first and/or last row of the tridiagonal axis,
built WITHOUT eliminating the ghost row: the grid is extended by the ghost
row(s), whose unknowns are kept, and each ghost row's equations are
replaced by the central-difference boundary equation
    lo: -(u_1 - u_{-1}) / 2h + alpha u_0 = g,
    hi:  (u_n - u_{n-2}) / 2h + alpha u_{n-1} = g.
The interior and boundary rows use the reference's 9/27-point stencils
(dense_ops.py) unchanged.  This is the independent check of the folded
band vectors and right-hand-side corrections of hh_robin2d / hh_robin3d."""
import numpy as np

from dense_ops import assemble2d, assemble3d


def _extend(coef, n, lo, hi):
    """Coefficient vectors over the extended rows; ghost rows get a copy of
    their neighbour's (their equations are overwritten afterwards)."""
    ext = []
    for v in coef:
        v = np.asarray(v, np.complex128).reshape(n)
        parts = ([v[:1]] if lo is not None else []) + [v] + ([v[-1:]] if hi is not None else [])
        ext.append(np.concatenate(parts))
    return ext


def _robin_rows(A, b, face, ne, h, lo, hi):
    """Overwrite the ghost rows of (A, b), layer-major with `face` unknowns per row."""
    def put(ghost, bnd, inner, alpha, g):
        for f in range(face):
            r = ghost * face + f
            A[r, :] = 0
            A[r, ghost * face + f] = 1.0 / (2 * h)
            A[r, bnd * face + f] = alpha
            A[r, inner * face + f] = -1.0 / (2 * h)
            b[r] = 0 if g is None else g.ravel()[f]
    if lo is not None:
        put(0, 1, 2, lo[0], lo[1])
    if hi is not None:
        put(ne - 1, ne - 2, ne - 3, hi[0], hi[1])


def solve_ghost2d(coef, Nx, Ny, h, F, lo=None, hi=None):
    """Interior solution (Ny, Nx) of the 2D problem with explicit ghost rows."""
    ext = _extend(coef, Ny, lo, hi)
    ne = len(ext[0])
    A = assemble2d(Nx, ne, ext)
    off = 1 if lo is not None else 0
    b = np.zeros(Nx * ne, np.complex128)
    b[off * Nx:(off + Ny) * Nx] = np.asarray(F).ravel()
    _robin_rows(A, b, Nx, ne, h, lo, hi)
    u = np.linalg.solve(A, b).reshape(ne, Nx)
    return u[off:off + Ny]


def solve_ghost3d(coef, Nx, Ny, Nz, h, F, lo=None, hi=None):
    """Interior solution (Nz, Ny, Nx) of the 3D problem with explicit ghost layers."""
    ext = _extend(coef, Nz, lo, hi)
    ne = len(ext[0])
    A = assemble3d(Nx, Ny, ne, np.stack(ext))
    off = 1 if lo is not None else 0
    face = Nx * Ny
    b = np.zeros(face * ne, np.complex128)
    b[off * face:(off + Nz) * face] = np.asarray(F).ravel()
    _robin_rows(A, b, face, ne, h, lo, hi)
    u = np.linalg.solve(A, b).reshape(ne, Ny, Nx)
    return u[off:off + Nz]


def folded_rhs(F, corr, n_face_rows):
    """F with the corrections added to its first / last row (layer)."""
    F = np.array(F, np.complex128, copy=True)
    if corr[0] is not None:
        F[0] += corr[0]
    if corr[1] is not None:
        F[n_face_rows - 1] += corr[1]
    return F
