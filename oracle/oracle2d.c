/*
 * 2D oracle (test infrastructure only; see oracle.h).
 * Restates 2D_Program/{main_functions,partial_derivatives,solver_functions}.
 */
#include "problems.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IX(i, j, Nx) ((long)(i) + (long)(j) * (long)(Nx))               /* ind   */
#define IE(i, j, Nx) ((long)(i) + 1 + ((long)(j) + 1) * ((long)(Nx) + 2)) /* ind_e */

void or2d_domain(int kind, const double *p, double dom[4]) {
    pr2 P;
    pr2_init(&P, kind, p);
    dom[0] = P.x0; dom[1] = P.x1; dom[2] = P.y0; dom[3] = P.y1;
}

/* coefficients.c:9-61 (second_order, fourth_order, sixth_order) */
static void coefficients2(int order, int Ny, const pr2 *P, double hx, double hy,
                          ocplx *bp, ocplx *ap, ocplx *b, ocplx *a, ocplx *bm, ocplx *am) {
    double hy2 = hy * hy, hy3 = hy2 * hy, hy4 = hy2 * hy2, Ryx = hy * hy / hx / hx;
    for (int j = 0; j < Ny; j++) {
        double y = P->y0 + (j + 1) * hy;
        if (order == OR_SIXTH) {                                   /* :44-61 */
            ocplx k2 = pr2_k2(P, y), k2p = pr2_k2(P, y + hy), k2m = pr2_k2(P, y - hy);
            ocplx ky2 = pr2_k2_y(P, y);
            ap[j] = 2.0 / 3.0 + hy2 * k2p / 90.0 + hy3 * ky2 * (hy2 * k2p / 6.0 + 2.0 / 3.0) / 20.0;
            bp[j] = 1.0 / 6.0 + hy2 * k2p / 90.0 + hy3 * ky2 / 120.0;
            a[j] = -10.0 / 3.0 + 41.0 * hy2 * k2 / 45.0 - hy4 * k2 * k2 / 20.0 + hy4 * pr2_k2_yy(P, y) / 20.0;
            b[j] = 2.0 / 3.0 + hy2 * k2 / 90.0;
            am[j] = 2.0 / 3.0 + hy2 * k2m / 90.0 - hy3 * ky2 * (hy2 * k2m / 6.0 + 2.0 / 3.0) / 20.0;
            bm[j] = 1.0 / 6.0 + hy2 * k2m / 90.0 - hy3 * ky2 / 120.0;
        } else if (order == OR_FOURTH) {                           /* :25-42 */
            ocplx k2 = pr2_k2(P, y), k2p = pr2_k2(P, y + hy), k2m = pr2_k2(P, y - hy);
            bp[j] = Ryx / 12.0 + 1.0 / 12.0;
            ap[j] = 5.0 / 6.0 - Ryx / 6.0 + hy2 * k2p / 12.0;
            b[j] = Ryx * (5.0 / 6.0) - 1.0 / 6.0 + hy2 * k2 / 12.0;
            a[j] = -Ryx * (5.0 / 3.0) - 5.0 / 3.0 + 2.0 * hy2 * k2 / 3.0;
            bm[j] = Ryx / 12.0 + 1.0 / 12.0;
            am[j] = 5.0 / 6.0 - Ryx / 6.0 + hy2 * k2m / 12.0;
        } else {                                                   /* :9-23 */
            ocplx kk = pr2_k(P, y);
            bp[j] = 0.0;
            ap[j] = 1.0;
            b[j] = Ryx;
            a[j] = -2.0 * Ryx - 2.0 + kk * kk * hy2;
            bm[j] = 0.0;
            am[j] = 1.0;
        }
    }
}

/* Constant-coefficient tridiagonal (1, diag, 1) solve along a line, as in
 * 4th_order_first_derivative.c:17-30 (diag 4) and
 * 4th_order_second_derivative.c:21-37 (diag 10). */
static void compact_solve(int n, double diag, ocplx *v, long stride, const double *U) {
    for (int j = 1; j < n; j++) v[j * stride] = v[j * stride] - (1.0 / U[j - 1]) * v[(j - 1) * stride];
    v[(n - 1) * stride] = v[(n - 1) * stride] / U[n - 1];
    for (int j = n - 2; j >= 0; j--) v[j * stride] = (v[j * stride] - v[(j + 1) * stride]) / U[j];
    (void)diag;
}

static double *compact_pivots(int n, double diag) {
    double *U = (double *)malloc(sizeof(double) * n);
    U[0] = diag;
    for (int j = 1; j < n; j++) U[j] = diag - 1.0 / U[j - 1];
    return U;
}

/* rhs.c:40-97 with the partial derivatives of 2D_Program/partial_derivatives */
static void sixth_order_rhs2(int Nx, int Ny, const pr2 *P, double hx, double hy,
                             ocplx *rhs, int threads) {
    long E = (long)(Nx + 2) * (Ny + 2), Nxy = (long)Nx * Ny;
    double hy2 = hy * hy, hy4 = hy2 * hy2, hx2 = hx * hx;
    ocplx *f_e = (ocplx *)malloc(sizeof(ocplx) * E);
    ocplx *fxx_e = (ocplx *)malloc(sizeof(ocplx) * E);
    ocplx *fyy_e = (ocplx *)malloc(sizeof(ocplx) * E);
    ocplx *fy = (ocplx *)malloc(sizeof(ocplx) * Nxy);
    ocplx *k2_e = (ocplx *)malloc(sizeof(ocplx) * (Ny + 2));
    ocplx *k2y = (ocplx *)malloc(sizeof(ocplx) * Ny);
    double *U4y = compact_pivots(Ny, 4.0), *U10y = compact_pivots(Ny, 10.0), *U10x = compact_pivots(Nx, 10.0);

    for (int j = 0; j < Ny + 2; j++) k2_e[j] = pr2_k2(P, P->y0 + j * hy);          /* :58 */
#pragma omp parallel for num_threads(threads)
    for (int j = 0; j < Ny + 2; j++)                                                  /* :60-66 */
        for (int i = 0; i < Nx + 2; i++) f_e[i + (long)j * (Nx + 2)] = pr2_f(P, P->x0 + i * hx, P->y0 + j * hy);

    /* k2_y_4th_order, 4th_order_first_derivative.c:4-34 */
    for (int j = 0; j < Ny; j++) k2y[j] = 3.0 * (k2_e[j + 2] - k2_e[j]) / hy;
    k2y[0] = k2y[0] - pr2_k2_y(P, P->y0);
    k2y[Ny - 1] = k2y[Ny - 1] - pr2_k2_y(P, P->y1);
    compact_solve(Ny, 4.0, k2y, 1, U4y);

    /* f_y_4th_order :36-74 and f_yy_4th_order 4th_order_second_derivative.c:64-122 */
#pragma omp parallel for num_threads(threads)
    for (int i = 0; i < Nx; i++) {
        double x = P->x0 + (i + 1) * hx;
        ocplx *col = (ocplx *)malloc(sizeof(ocplx) * Ny);
        for (int j = 0; j < Ny; j++) fy[IX(i, j, Nx)] = 3.0 * (f_e[IE(i, j + 1, Nx)] - f_e[IE(i, j - 1, Nx)]) / hy;
        fy[IX(i, 0, Nx)] -= pr2_f_y(P, x, P->y0);
        fy[IX(i, Ny - 1, Nx)] -= pr2_f_y(P, x, P->y1);
        compact_solve(Ny, 4.0, fy + i, Nx, U4y);
        for (int j = 0; j < Ny; j++)
            col[j] = 12.0 * (f_e[IE(i, j + 1, Nx)] - 2.0 * f_e[IE(i, j, Nx)] + f_e[IE(i, j - 1, Nx)]) / hy / hy;
        col[0] -= pr2_f_yy(P, x, P->y0);
        col[Ny - 1] -= pr2_f_yy(P, x, P->y1);
        compact_solve(Ny, 10.0, col, 1, U10y);
        for (int j = 0; j < Ny; j++) fyy_e[IE(i, j, Nx)] = col[j];
        free(col);
    }
    /* f_xx_4th_order 4th_order_second_derivative.c:4-62 (pivot index fixed
     * to Nx-1, see line 34 of that file) */
#pragma omp parallel for num_threads(threads)
    for (int j = 0; j < Ny; j++) {
        double y = P->y0 + (j + 1) * hy;
        ocplx *row = (ocplx *)malloc(sizeof(ocplx) * Nx);
        for (int i = 0; i < Nx; i++)
            row[i] = 12.0 * (f_e[IE(i + 1, j, Nx)] - 2.0 * f_e[IE(i, j, Nx)] + f_e[IE(i - 1, j, Nx)]) / hx / hx;
        row[0] -= pr2_f_xx(P, P->x0, y);
        row[Nx - 1] -= pr2_f_xx(P, P->x1, y);
        compact_solve(Nx, 10.0, row, 1, U10x);
        for (int i = 0; i < Nx; i++) fxx_e[IE(i, j, Nx)] = row[i];
        free(row);
    }
    /* ghost ring of fxx_e and fyy_e: analytic values (:41-51, :101-111) */
    for (int i = 0; i < Nx + 2; i++) {
        double x = P->x0 + i * hx;
        fxx_e[i] = pr2_f_xx(P, x, P->y0);
        fxx_e[i + (long)(Ny + 1) * (Nx + 2)] = pr2_f_xx(P, x, P->y1);
        fyy_e[i] = pr2_f_yy(P, x, P->y0);
        fyy_e[i + (long)(Ny + 1) * (Nx + 2)] = pr2_f_yy(P, x, P->y1);
    }
    for (int j = 0; j < Ny + 2; j++) {
        double y = P->y0 + j * hy;
        fxx_e[(long)j * (Nx + 2)] = pr2_f_xx(P, P->x0, y);
        fxx_e[Nx + 1 + (long)j * (Nx + 2)] = pr2_f_xx(P, P->x1, y);
        fyy_e[(long)j * (Nx + 2)] = pr2_f_yy(P, P->x0, y);
        fyy_e[Nx + 1 + (long)j * (Nx + 2)] = pr2_f_yy(P, P->x1, y);
    }
    /* derivative_{xx,yy}_2nd_order (2nd_order.c:4-24) and the combination :78-86 */
#pragma omp parallel for num_threads(threads)
    for (int j = 0; j < Ny; j++) {
        double y = P->y0 + (j + 1) * hy;
        ocplx kk = pr2_k(P, y), k2c = kk * kk;
        for (int i = 0; i < Nx; i++) {
            ocplx fxxxx = (fxx_e[IE(i - 1, j, Nx)] - 2.0 * fxx_e[IE(i, j, Nx)] + fxx_e[IE(i + 1, j, Nx)]) / hx2;
            ocplx fxxyy = (fxx_e[IE(i, j - 1, Nx)] - 2.0 * fxx_e[IE(i, j, Nx)] + fxx_e[IE(i, j + 1, Nx)]) / hy2;
            ocplx fyyyy = (fyy_e[IE(i, j - 1, Nx)] - 2.0 * fyy_e[IE(i, j, Nx)] + fyy_e[IE(i, j + 1, Nx)]) / hy2;
            rhs[IX(i, j, Nx)] = hy2 * ((1.0 - k2c * hy2 / 20.0) * f_e[IE(i, j, Nx)] +
                                       hy2 * (fxx_e[IE(i, j, Nx)] + fyy_e[IE(i, j, Nx)]) / 12.0 +
                                       hy4 * (fxxxx + fyyyy) / 360.0 + hy4 * fxxyy / 90.0 +
                                       hy4 * k2y[j] * fy[IX(i, j, Nx)] / 60.0);
        }
    }
    free(f_e); free(fxx_e); free(fyy_e); free(fy); free(k2_e); free(k2y);
    free(U4y); free(U10y); free(U10x);
}

int or2d_setup(int order, int Nx, int Ny, int kind, const double *p,
               ocplx *bp, ocplx *ap, ocplx *b, ocplx *a, ocplx *bm, ocplx *am,
               ocplx *rhs, ocplx *sol_analytic, int threads) {
    if (Nx < 2 || Ny < 2) return -1;
    if (order == OR_SIXTH && Nx != Ny) return -2;   /* setup.c:43-46 */
    if (threads < 1) threads = 1;
    pr2 P;
    pr2_init(&P, kind, p);
    double hx = (P.x1 - P.x0) / (Nx + 1), hy = (P.y1 - P.y0) / (Ny + 1);  /* setup.c:48-49 */
    coefficients2(order, Ny, &P, hx, hy, bp, ap, b, a, bm, am);

    /* rhs.c:100-124 */
#pragma omp parallel for num_threads(threads)
    for (int j = 0; j < Ny; j++) {
        double y = P.y0 + (j + 1) * hy;
        for (int i = 0; i < Nx; i++) {
            double x = P.x0 + (i + 1) * hx;
            sol_analytic[IX(i, j, Nx)] = pr2_u(&P, x, y);
            if (order == OR_FOURTH)                                   /* :21-38 */
                rhs[IX(i, j, Nx)] = hy * hy * ((2.0 / 3.0) * pr2_f(&P, x, y) +
                                    (1.0 / 12.0) * (pr2_f(&P, x, y - hy) + pr2_f(&P, x, y + hy) +
                                                    pr2_f(&P, x - hx, y) + pr2_f(&P, x + hx, y)));
            else if (order != OR_SIXTH)                               /* :4-19 */
                rhs[IX(i, j, Nx)] = hy * hy * pr2_f(&P, x, y);
        }
    }
    if (order == OR_SIXTH) sixth_order_rhs2(Nx, Ny, &P, hx, hy, rhs, threads);

    /* Dirichlet closure rhs.c:126-148: subtract the stencil weight times the
     * boundary value for every stencil neighbour that lies on the boundary. */
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            if (i > 0 && i < Nx - 1 && j > 0 && j < Ny - 1) continue;
            ocplx acc = 0;
            for (int dj = -1; dj <= 1; dj++) {
                for (int di = -1; di <= 1; di++) {
                    int ii = i + di, jj = j + dj;
                    if (ii >= 0 && ii < Nx && jj >= 0 && jj < Ny) continue;
                    ocplx w = dj == 1 ? (di == 0 ? ap[j] : bp[j])
                            : dj == 0 ? (di == 0 ? a[j] : b[j])
                                      : (di == 0 ? am[j] : bm[j]);
                    acc += w * pr2_u(&P, P.x0 + (ii + 1) * hx, P.y0 + (jj + 1) * hy);
                }
            }
            rhs[IX(i, j, Nx)] -= acc;
        }
    }
    return 0;
}

/* LU.c:3-46 */
void or2d_LU(int Nx, int Ny, const ocplx *bp, const ocplx *ap, const ocplx *b,
             const ocplx *a, const ocplx *bm, const ocplx *am,
             ocplx *L, ocplx *U, ocplx *Up, int threads) {
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads)
    for (int i = 0; i < Nx; i++) {
        double lam = 2.0 * cos((i + 1) * M_PI / (Nx + 1));          /* :12-14 */
        long my = (long)i * Ny;
        L[my] = 0.0;
        U[my] = b[0] * lam + a[0];
        Up[my] = bp[0] * lam + ap[0];
        for (int j = 1; j < Ny; j++) {                               /* :32-40 */
            L[j + my] = (bm[j] * lam + am[j]) / U[j + my - 1];
            U[j + my] = (b[j] * lam + a[j]) - L[j + my] * Up[j + my - 1];
            Up[j + my] = bp[j] * lam + ap[j];
        }
    }
}

/* solver.c:14-117 + DST.c:6-44, with the DST plan created inside the call
 * as the reference does (:56-64) */
void or2d_solver(int Nx, int Ny, const ocplx *rhs, const ocplx *L,
                 const ocplx *U, const ocplx *Up, ocplx *sol, int threads) {
    or_dstplan *plan = or_dst_plan(Nx);                              /* :56-64 */
    or2d_solver_planned(plan, Nx, Ny, rhs, L, U, Up, sol, threads);
    or_dst_plan_free(plan);
}

/* the same solve with a caller-held plan (steady state) */
void or2d_solver_planned(const or_dstplan *plan, int Nx, int Ny, const ocplx *rhs, const ocplx *L,
                         const ocplx *U, const ocplx *Up, ocplx *sol, int threads) {
    if (threads < 1) threads = 1;
    long Nxy = (long)Nx * Ny;
    ocplx *rhat = (ocplx *)malloc(sizeof(ocplx) * Nxy);
    ocplx *xhat = (ocplx *)malloc(sizeof(ocplx) * Nxy);
    or_dst_lines_plan(plan, Ny, rhs, Nx, 1, rhat, Nx, 1, threads);   /* forwardDST */
#pragma omp parallel for num_threads(threads)
    for (int i = 0; i < Nx; i++) {                                   /* :71-86 */
        ocplx *y = (ocplx *)malloc(sizeof(ocplx) * Ny);
        long mx = (long)i * Ny;
        y[0] = rhat[i];
        for (int j = 1; j < Ny; j++) y[j] = rhat[IX(i, j, Nx)] - L[j + mx] * y[j - 1];
        xhat[Ny - 1 + mx] = y[Ny - 1] / U[Ny - 1 + mx];
        for (int j = Ny - 2; j >= 0; j--) xhat[j + mx] = (y[j] - Up[j + mx] * xhat[j + 1 + mx]) / U[j + mx];
        free(y);
    }
    /* reverseDST: row j gathers xhat[j + i*Ny] (stride Ny) */
    or_dst_lines_plan(plan, Ny, xhat, 1, Ny, sol, Nx, 1, threads);
    free(rhat);
    free(xhat);
}

/* residual.c:3-34 */
void or2d_residual(int Nx, int Ny, const ocplx *bp, const ocplx *ap,
                   const ocplx *b, const ocplx *a, const ocplx *bm,
                   const ocplx *am, const ocplx *sol, const ocplx *rhs,
                   ocplx *res) {
    long E = (long)(Nx + 2) * (Ny + 2);
    ocplx *s = (ocplx *)calloc(E, sizeof(ocplx));
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) s[IE(i, j, Nx)] = sol[IX(i, j, Nx)];
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            ocplx top = bp[j] * s[IE(i - 1, j + 1, Nx)] + ap[j] * s[IE(i, j + 1, Nx)] + bp[j] * s[IE(i + 1, j + 1, Nx)];
            ocplx mid = b[j] * s[IE(i - 1, j, Nx)] + a[j] * s[IE(i, j, Nx)] + b[j] * s[IE(i + 1, j, Nx)];
            ocplx bot = bm[j] * s[IE(i - 1, j - 1, Nx)] + am[j] * s[IE(i, j - 1, Nx)] + bm[j] * s[IE(i + 1, j - 1, Nx)];
            res[IX(i, j, Nx)] = top + mid + bot - rhs[IX(i, j, Nx)];
        }
    }
    free(s);
}

/* Inputs of the GPU right-hand-side assembly for the same problem:
 * f_e  : f at every point of the extended (Nx+2)x(Ny+2) grid
 * u_e  : u (Dirichlet data) at every point of the extended grid
 * fy_b : f_y at (x_i, y0) [0..Nx) and (x_i, y1) [Nx..2Nx)
 * fxx_b, fyy_b : ghost ring of f_xx / f_yy, ordered bottom row (Nx+2),
 *        top row (Nx+2), left column (Ny+2), right column (Ny+2)
 * k2_e : k^2 at y0 + j*hy, j = 0..Ny+1
 * k2y_b: (k^2)_y at y0 and y1 */
void or2d_fields(int Nx, int Ny, int kind, const double *p, ocplx *f_e,
                 ocplx *u_e, ocplx *fy_b, ocplx *fxx_b, ocplx *fyy_b,
                 ocplx *k2_e, ocplx *k2y_b) {
    pr2 P;
    pr2_init(&P, kind, p);
    double hx = (P.x1 - P.x0) / (Nx + 1), hy = (P.y1 - P.y0) / (Ny + 1);
    for (int j = 0; j < Ny + 2; j++) {
        for (int i = 0; i < Nx + 2; i++) {
            double x = P.x0 + i * hx, y = P.y0 + j * hy;
            f_e[i + (long)j * (Nx + 2)] = pr2_f(&P, x, y);
            u_e[i + (long)j * (Nx + 2)] = pr2_u(&P, x, y);
        }
        k2_e[j] = pr2_k2(&P, P.y0 + j * hy);
    }
    for (int i = 0; i < Nx; i++) {
        double x = P.x0 + (i + 1) * hx;
        fy_b[i] = pr2_f_y(&P, x, P.y0);
        fy_b[Nx + i] = pr2_f_y(&P, x, P.y1);
    }
    int X = Nx + 2, Y = Ny + 2;
    for (int i = 0; i < X; i++) {
        double x = P.x0 + i * hx;
        fxx_b[i] = pr2_f_xx(&P, x, P.y0);
        fxx_b[X + i] = pr2_f_xx(&P, x, P.y1);
        fyy_b[i] = pr2_f_yy(&P, x, P.y0);
        fyy_b[X + i] = pr2_f_yy(&P, x, P.y1);
    }
    for (int j = 0; j < Y; j++) {
        double y = P.y0 + j * hy;
        fxx_b[2 * X + j] = pr2_f_xx(&P, P.x0, y);
        fxx_b[2 * X + Y + j] = pr2_f_xx(&P, P.x1, y);
        fyy_b[2 * X + j] = pr2_f_yy(&P, P.x0, y);
        fyy_b[2 * X + Y + j] = pr2_f_yy(&P, P.x1, y);
    }
    k2y_b[0] = pr2_k2_y(&P, P.y0);
    k2y_b[1] = pr2_k2_y(&P, P.y1);
}
