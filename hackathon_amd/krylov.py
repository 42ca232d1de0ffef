"""Krylov mode: the direct solver as the preconditioner of a restarted GMRES
for operators it does not diagonalise itself.

The DST diagonalisation needs coefficients that vary along the tridiagonal
axis only (y in 2D).  A wave number that also varies along x, k(x, y),
breaks it; the direct solver for the x-averaged k^2 is then a close
approximation of the operator, and right-preconditioned GMRES converges
in a handful of iterations (the reference's paper cites this use of
the direct solver, as a preconditioner for nonconstant-coefficient and
Sommerfeld-like problems: Resources/paper.pdf p.1, l.27-41).

The operator apply and the preconditioner are the library's HIP kernels
(the compact residual kernel plus a diagonal term, and the direct solve);
the Krylov vector algebra (dot products, updates of a few vectors) runs in
PyTorch on the same device and stream.
"""
import numpy as np

from . import helmholtz as H


def gmres(apply_a, b, precond=None, tol=1e-10, restart=30, maxiter=300):
    """Solve A x = b with right-preconditioned restarted GMRES(restart).

    apply_a(v) -> A v and precond(v) -> M^{-1} v take and return complex128
    GPU tensors of b's shape.  Stops when ||b - A x||_2 <= tol ||b||_2 (the
    true residual, recomputed at every restart) or after maxiter inner
    iterations.  The Krylov basis is one device tensor; each new vector is
    orthogonalised against it by classical Gram-Schmidt applied twice (two
    matrix-vector products on the device), so an iteration reads back one
    Hessenberg column and one norm to the host, not one inner product per
    basis vector.  Returns (x, info) with info = {iterations,
    rel_residual, converged}."""
    torch = H._torch()
    M = precond if precond is not None else (lambda v: v)
    shape = b.shape
    n = b.numel()
    nb = float(torch.linalg.vector_norm(b))
    x = torch.zeros_like(b)
    if nb == 0.0:
        return x, {"iterations": 0, "rel_residual": 0.0, "converged": True}
    V = torch.empty((restart + 1, n), dtype=b.dtype, device=b.device)
    r = b.clone()
    rel = 1.0
    it = 0
    while True:
        beta = float(torch.linalg.vector_norm(r))
        rel = beta / nb
        if rel <= tol or it >= maxiter:
            break
        V[0] = (r / beta).reshape(-1)
        Hm = np.zeros((restart + 1, restart), np.complex128)
        cs = np.zeros(restart, np.complex128)
        sn = np.zeros(restart, np.complex128)
        g = np.zeros(restart + 1, np.complex128)
        g[0] = beta
        k = 0
        for j in range(restart):
            w = apply_a(M(V[j].reshape(shape))).reshape(-1)
            B = V[: j + 1]
            h1 = torch.mv(B.conj(), w)           # CGS, then once more for orthogonality
            w = w - torch.mv(B.transpose(0, 1), h1)
            h2 = torch.mv(B.conj(), w)
            w = w - torch.mv(B.transpose(0, 1), h2)
            hn_t = torch.linalg.vector_norm(w)
            col = torch.cat([h1 + h2, hn_t.to(b.dtype).reshape(1)]).cpu().numpy()   # the one read-back
            Hm[: j + 1, j] = col[: j + 1]
            hn = float(col[j + 1].real)
            Hm[j + 1, j] = hn
            for i in range(j):       # earlier Givens rotations on the new column
                t = cs[i] * Hm[i, j] + sn[i] * Hm[i + 1, j]
                Hm[i + 1, j] = -np.conj(sn[i]) * Hm[i, j] + cs[i] * Hm[i + 1, j]
                Hm[i, j] = t
            # rotation [[c, s], [-conj(s), c]] (c real) zeroing the subdiagonal
            a = Hm[j, j]
            den = np.hypot(abs(a), hn)
            if abs(a) == 0.0:
                cs[j], sn[j] = 0.0, 1.0
            else:
                cs[j] = abs(a) / den
                sn[j] = (a / abs(a)) * hn / den
            Hm[j, j] = cs[j] * a + sn[j] * hn
            Hm[j + 1, j] = 0.0
            g[j + 1] = -np.conj(sn[j]) * g[j]
            g[j] = cs[j] * g[j]
            it += 1
            k = j + 1
            if hn == 0.0 or abs(g[j + 1]) / nb <= tol or it >= maxiter:
                break
            V[j + 1] = w / hn
        y = np.linalg.solve(np.triu(Hm[:k, :k]), g[:k])
        yt = torch.from_numpy(y).to(device=b.device, dtype=b.dtype)
        upd = torch.mv(V[:k].transpose(0, 1), yt).reshape(shape)
        x = x + M(upd)
        r = b - apply_a(x)
    return x, {"iterations": it, "rel_residual": rel, "converged": rel <= tol}


class VariableK2D:
    """Second-order five-point Helmholtz operator on the unit square with a
    wave number varying in x and y (homogeneous Dirichlet boundary):

        (A u)_ij = u_{i-1,j} + u_{i+1,j} + u_{i,j-1} + u_{i,j+1}
                   + (h^2 k^2(x_i, y_j) - 4) u_ij

    (the reference's second-order stencil, coefficients.c:9-23, with k^2
    taken per point).  k2 is an (N, N) array of k^2 at the interior points,
    [j][i] = k^2(x_i, y_j).  The preconditioner is the direct solve of the
    same stencil with k^2 averaged along x; A u is its residual kernel plus
    the diagonal h^2 (k^2 - mean_x k^2) u."""

    def __init__(self, ctx, N, k2):
        torch = H._torch()
        k2 = np.asarray(k2, dtype=np.complex128).reshape(N, N)
        self.N, self.h = N, 1.0 / (N + 1)
        kbar = k2.mean(axis=1)                                   # per row j
        k2e = np.concatenate([[kbar[0]], kbar, [kbar[-1]]])      # extended rows (only interior rows enter)
        coef = H.coefficients2d(2, N, N, self.h, self.h, k2e)
        self.plan = H.Plan2D(ctx, N, N, coef)
        d = self.h ** 2 * (k2 - kbar[:, None])
        self.diag = torch.from_numpy(np.ascontiguousarray(d)).to("cuda")
        self.zero = torch.zeros((N, N), dtype=torch.complex128, device="cuda")

    def apply(self, u):
        u = u.reshape(self.N, self.N)
        return self.plan.residual(u, self.zero) + self.diag * u

    def precond(self, v):
        return self.plan.solve(v.reshape(self.N, self.N).contiguous())

    def solve(self, F, tol=1e-12, restart=30, maxiter=300):
        """x with A x = F (F: (N, N) complex128 GPU tensor)."""
        return gmres(self.apply, F.reshape(self.N, self.N).contiguous(), self.precond, tol, restart, maxiter)


class VariableK3D:
    """Second-order seven-point Helmholtz operator on the unit cube with a
    wave number varying in x, y and z (homogeneous Dirichlet boundary):

        (A u)_ijl = u_{i-1} + u_{i+1} + u_{j-1} + u_{j+1} + u_{l-1} + u_{l+1}
                    + (h^2 k^2(x_i, y_j, z_l) - 6) u_ijl

    (the reference's 3D second-order stencil, 3D_Program/main_functions/
    coefficients.c:7-32, with h = hx = hy = hz and k^2 taken per point).  k2
    is an (N, N, N) array of k^2 at the interior points, [l][j][i] =
    k^2(x_i, y_j, z_l).  The preconditioner is the direct solve of the same
    stencil with k^2 averaged over x and y in each z layer (what the 3D DST
    diagonalisation allows); A u is its residual kernel plus the diagonal
    h^2 (k^2 - mean_xy k^2) u."""

    def __init__(self, ctx, N, k2):
        torch = H._torch()
        k2 = np.asarray(k2, dtype=np.complex128).reshape(N, N, N)
        self.N, self.h = N, 1.0 / (N + 1)
        kbar = k2.mean(axis=(1, 2))                              # per layer l
        k2e = np.concatenate([[kbar[0]], kbar, [kbar[-1]]])
        coef = H.coefficients3d(2, N, N, N, self.h, self.h, self.h, k2e)
        self.plan = H.Plan3D(ctx, N, N, N, coef)
        d = self.h ** 2 * (k2 - kbar[:, None, None])
        self.diag = torch.from_numpy(np.ascontiguousarray(d)).to("cuda")
        self.zero = torch.zeros((N, N, N), dtype=torch.complex128, device="cuda")

    def apply(self, u):
        u = u.reshape(self.N, self.N, self.N)
        return self.plan.residual(u, self.zero) + self.diag * u

    def precond(self, v):
        return self.plan.solve(v.reshape(self.N, self.N, self.N).contiguous())

    def solve(self, F, tol=1e-12, restart=30, maxiter=300):
        """x with A x = F (F: (N, N, N) complex128 GPU tensor)."""
        return gmres(self.apply, F.reshape(self.N, self.N, self.N).contiguous(), self.precond, tol, restart, maxiter)
