/*
 * 3D oracle (test infrastructure only; see oracle.h).
 * Restates 3D_Program/{main_functions,partial_derivatives} and the CPU
 * branch of 3D_Program/solver_functions/solver.c:263-413 with DST.c:148-204.
 *
 * Reference defects not reproduced (the restatement follows the intended
 * arithmetic):
 *  - rhs.c:149 declares the five 6th-order RHS terms `double`, dropping the
 *    imaginary part of complex data; here they stay complex.
 *  - 4th_order_partials.c:143-145 writes the x = x1 ghost face of fxx to a
 *    wrong index and leaves the y/z ghost faces of fxx, fyy, fzz unset
 *    (zero pages in practice); here the reference faces come from the
 *    analytic second derivatives and the other ghost faces are zero.
 */
#include "problems.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define I3(i, j, l) ((long)(i) + (long)(j) * Nx + (long)(l) * Nxy)
#define E3(i, j, l) ((long)(i) + (long)(j) * X + (long)(l) * XY)  /* extended, 0..N+1 */

void or3d_domain(int kind, const double *p, double dom[6]) {
    pr3 P;
    pr3_init(&P, kind, p);
    dom[0] = P.x0; dom[1] = P.x1; dom[2] = P.y0; dom[3] = P.y1; dom[4] = P.z0; dom[5] = P.z1;
}

static double *pivots(int n, double diag) {
    double *U = (double *)malloc(sizeof(double) * n);
    U[0] = diag;
    for (int j = 1; j < n; j++) U[j] = diag - 1.0 / U[j - 1];
    return U;
}

/* (1, diag, 1) solve, 4th_order_partials.c:118-139 */
static void csolve(int n, ocplx *v, long s, const double *U) {
    for (int j = 1; j < n; j++) v[j * s] = v[j * s] - (1.0 / U[j - 1]) * v[(j - 1) * s];
    v[(n - 1) * s] = v[(n - 1) * s] / U[n - 1];
    for (int j = n - 2; j >= 0; j--) v[j * s] = (v[j * s] - v[(j + 1) * s]) / U[j];
}

/* derivative_kz_4th_order / derivative_kzz_4th_order, 4th_order_partials.c:296-385 */
static void k2_derivs(int Nz, double hz, const pr3 *P, const ocplx *k2e, ocplx *kz, ocplx *kzz) {
    double *U4 = pivots(Nz, 4.0), *U10 = pivots(Nz, 10.0);
    for (int l = 0; l < Nz; l++) {
        kz[l] = 3.0 * (k2e[l + 2] - k2e[l]) / hz;
        kzz[l] = 12.0 * (k2e[l] - 2.0 * k2e[l + 1] + k2e[l + 2]) / hz / hz;
    }
    kz[0] -= pr3_k2_z(P, P->z0);
    kz[Nz - 1] -= pr3_k2_z(P, P->z1);
    kzz[0] -= pr3_k2_zz(P, P->z0);
    kzz[Nz - 1] -= pr3_k2_zz(P, P->z1);
    csolve(Nz, kz, 1, U4);
    csolve(Nz, kzz, 1, U10);
    free(U4); free(U10);
}

/* coefficients.c:7-111; coef = 12 arrays of Nz: a,b,c,d,ap,bp,cp,dp,am,bm,cm,dm */
static void coefficients3(int order, int Nz, const pr3 *P, double hx, double hy, double hz,
                          const ocplx *k2, ocplx *coef) {
    ocplx *a = coef, *b = coef + Nz, *c = coef + 2 * Nz, *d = coef + 3 * Nz;
    ocplx *ap = coef + 4 * Nz, *bp = coef + 5 * Nz, *cp = coef + 6 * Nz, *dp = coef + 7 * Nz;
    ocplx *am = coef + 8 * Nz, *bm = coef + 9 * Nz, *cm = coef + 10 * Nz, *dm = coef + 11 * Nz;
    double Rzx = hz * hz / hx / hx, Rzy = hz * hz / hy / hy, h2 = hz * hz;
    if (order == OR_SIXTH) {                                                  /* :67-103 */
        ocplx *kp = (ocplx *)malloc(sizeof(ocplx) * Nz), *kpp = (ocplx *)malloc(sizeof(ocplx) * Nz);
        k2_derivs(Nz, hz, P, k2, kp, kpp);
        for (int i = 0; i < Nz; i++) {
            a[i] = 1.0 / 10.0 + h2 * k2[i + 1] / 90.0;
            b[i] = 7.0 / 15.0 - h2 * k2[i + 1] / 90.0;
            c[i] = 7.0 / 15.0 - h2 * k2[i + 1] / 90.0;
            d[i] = -64.0 / 15.0 + 14.0 * h2 * k2[i + 1] / 15.0 - h2 * h2 * k2[i + 1] * k2[i + 1] / 20.0 +
                   h2 * h2 * kpp[i] / 20.0;
            ap[i] = 1.0 / 30.0;
            bp[i] = 1.0 / 10.0 + h2 * k2[i + 2] / 90.0 + (kp[i] * h2 * hz / 120.0);
            cp[i] = 1.0 / 10.0 + h2 * k2[i + 2] / 90.0 + (kp[i] * h2 * hz / 120.0);
            dp[i] = 7.0 / 15.0 - h2 * k2[i + 2] / 90.0 + (kp[i] * h2 * hz / 20.0) * (1.0 / 3.0 + k2[i + 2] * h2 / 6.0);
            am[i] = 1.0 / 30.0;
            bm[i] = 1.0 / 10.0 + h2 * k2[i] / 90.0 - (kp[i] * h2 * hz / 120.0);
            cm[i] = 1.0 / 10.0 + h2 * k2[i] / 90.0 - (kp[i] * h2 * hz / 120.0);
            dm[i] = 7.0 / 15.0 - h2 * k2[i] / 90.0 - (kp[i] * h2 * hz / 20.0) * (1.0 / 3.0 + k2[i] * h2 / 6.0);
        }
        free(kp); free(kpp);
    } else if (order == OR_FOURTH) {                                          /* :35-65 */
        for (int i = 0; i < Nz; i++) {
            a[i] = (1.0 / 12.0) * (Rzx + Rzy);
            b[i] = (1.0 / 6.0) * (4.0 * Rzx - Rzy - 1.0 + 0.5 * h2 * k2[i + 1]);
            c[i] = (1.0 / 6.0) * (4.0 * Rzy - Rzx - 1.0 + 0.5 * h2 * k2[i + 1]);
            d[i] = -4.0 * (1.0 + Rzx + Rzy) / 3.0 + 0.5 * h2 * k2[i + 1];
            ap[i] = 0.0;
            bp[i] = (1.0 / 12.0) * (1.0 + Rzx);
            cp[i] = (1.0 / 12.0) * (1.0 + Rzy);
            dp[i] = 1.0 - (1.0 / 6.0) * (2.0 + Rzx + Rzy) + (1.0 / 12.0) * h2 * k2[i + 2];
            am[i] = 0.0;
            bm[i] = (1.0 / 12.0) * (1.0 + Rzx);
            cm[i] = (1.0 / 12.0) * (1.0 + Rzy);
            dm[i] = 1.0 - (1.0 / 6.0) * (2.0 + Rzx + Rzy) + (1.0 / 12.0) * h2 * k2[i];
        }
    } else {                                                                  /* :7-33 */
        for (int i = 0; i < Nz; i++) {
            a[i] = 0; b[i] = Rzx; c[i] = Rzy;
            d[i] = -2.0 * (Rzx + Rzy + 1.0) + h2 * k2[i + 1];
            ap[i] = 0; bp[i] = 0; cp[i] = 0; dp[i] = 1;
            am[i] = 0; bm[i] = 0; cm[i] = 0; dm[i] = 1;
        }
    }
}

int or3d_setup(int order, int Nx, int Ny, int Nz, int kind, const double *p,
               ocplx *coef, ocplx *rhs, ocplx *sol_analytic, int threads) {
    if (Nx < 2 || Ny < 2 || Nz < 2) return -1;
    if (threads < 1) threads = 1;
    pr3 P;
    pr3_init(&P, kind, p);
    long Nxy = (long)Nx * Ny, X = Nx + 2, XY = (long)(Nx + 2) * (Ny + 2);
    long E = XY * (Nz + 2);
    double hx = (P.x1 - P.x0) / (Nx + 1), hy = (P.y1 - P.y0) / (Ny + 1), hz = (P.z1 - P.z0) / (Nz + 1);
    double h2 = hz * hz;
    ocplx *k2 = (ocplx *)malloc(sizeof(ocplx) * (Nz + 2));
    for (int l = 0; l < Nz + 2; l++) k2[l] = pr3_k2(&P, P.z0 + hz * l);    /* background3D */
    coefficients3(order, Nz, &P, hx, hy, hz, k2, coef);
    ocplx *ap = coef + 4 * Nz, *bp = coef + 5 * Nz, *cp = coef + 6 * Nz, *dp = coef + 7 * Nz;
    ocplx *am = coef + 8 * Nz, *bm = coef + 9 * Nz, *cm = coef + 10 * Nz, *dm = coef + 11 * Nz;

    /* rhs.c:229-267: UE_ext and rhs_ext = hz^2 f on the extended grid */
    ocplx *UE = (ocplx *)malloc(sizeof(ocplx) * E);
    ocplx *F = (ocplx *)malloc(sizeof(ocplx) * E);
#pragma omp parallel for num_threads(threads)
    for (int l = 0; l < Nz + 2; l++)
        for (int j = 0; j < Ny + 2; j++)
            for (int i = 0; i < Nx + 2; i++) {
                double x = P.x0 + i * hx, y = P.y0 + j * hy, z = P.z0 + l * hz;
                UE[E3(i, j, l)] = pr3_u(&P, x, y, z);
                F[E3(i, j, l)] = h2 * pr3_f(&P, x, y, z);
            }
    for (int l = 0; l < Nz; l++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) sol_analytic[I3(i, j, l)] = UE[E3(i + 1, j + 1, l + 1)];

    if (order == OR_SIXTH) {                                                  /* rhs.c:67-194 */
        ocplx *fxx = (ocplx *)calloc(E, sizeof(ocplx));
        ocplx *fyy = (ocplx *)calloc(E, sizeof(ocplx));
        ocplx *fzz = (ocplx *)calloc(E, sizeof(ocplx));
        double *Ux = pivots(Nx, 10.0), *Uy = pivots(Ny, 10.0), *Uz = pivots(Nz, 10.0);
        /* reference ghost faces (truef{xx,yy,zz}_ext, rhs.c:90-109), analytic */
        for (int l = 0; l < Nz + 2; l++)
            for (int j = 0; j < Ny + 2; j++) {
                double y = P.y0 + j * hy, z = P.z0 + l * hz;
                fxx[E3(0, j, l)] = h2 * pr3_f_xx(&P, P.x0, y, z);
                fxx[E3(Nx + 1, j, l)] = h2 * pr3_f_xx(&P, P.x1, y, z);
            }
        for (int l = 0; l < Nz + 2; l++)
            for (int i = 0; i < Nx + 2; i++) {
                double x = P.x0 + i * hx, z = P.z0 + l * hz;
                fyy[E3(i, 0, l)] = h2 * pr3_f_yy(&P, x, P.y0, z);
                fyy[E3(i, Ny + 1, l)] = h2 * pr3_f_yy(&P, x, P.y1, z);
            }
        for (int j = 0; j < Ny + 2; j++)
            for (int i = 0; i < Nx + 2; i++) {
                double x = P.x0 + i * hx, y = P.y0 + j * hy;
                fzz[E3(i, j, 0)] = h2 * pr3_f_zz(&P, x, y, P.z0);
                fzz[E3(i, j, Nz + 1)] = h2 * pr3_f_zz(&P, x, y, P.z1);
            }
        /* derivative_xx_4th_order :94-160 */
#pragma omp parallel for num_threads(threads)
        for (int l = 0; l < Nz; l++) {
            ocplx *v = (ocplx *)malloc(sizeof(ocplx) * Nx);
            for (int j = 0; j < Ny; j++) {
                for (int i = 0; i < Nx; i++)
                    v[i] = 12.0 * (F[E3(i, j + 1, l + 1)] - 2.0 * F[E3(i + 1, j + 1, l + 1)] + F[E3(i + 2, j + 1, l + 1)]) / hx / hx;
                v[0] -= fxx[E3(0, j + 1, l + 1)];
                v[Nx - 1] -= fxx[E3(Nx + 1, j + 1, l + 1)];
                csolve(Nx, v, 1, Ux);
                for (int i = 0; i < Nx; i++) fxx[E3(i + 1, j + 1, l + 1)] = v[i];
            }
            free(v);
        }
        /* derivative_yy_4th_order :162-227 */
#pragma omp parallel for num_threads(threads)
        for (int l = 0; l < Nz; l++) {
            ocplx *v = (ocplx *)malloc(sizeof(ocplx) * Ny);
            for (int i = 0; i < Nx; i++) {
                for (int j = 0; j < Ny; j++)
                    v[j] = 12.0 * (F[E3(i + 1, j, l + 1)] - 2.0 * F[E3(i + 1, j + 1, l + 1)] + F[E3(i + 1, j + 2, l + 1)]) / hy / hy;
                v[0] -= fyy[E3(i + 1, 0, l + 1)];
                v[Ny - 1] -= fyy[E3(i + 1, Ny + 1, l + 1)];
                csolve(Ny, v, 1, Uy);
                for (int j = 0; j < Ny; j++) fyy[E3(i + 1, j + 1, l + 1)] = v[j];
            }
            free(v);
        }
        /* derivative_zz_4th_order :229-294 */
#pragma omp parallel for num_threads(threads)
        for (int j = 0; j < Ny; j++) {
            ocplx *v = (ocplx *)malloc(sizeof(ocplx) * Nz);
            for (int i = 0; i < Nx; i++) {
                for (int l = 0; l < Nz; l++)
                    v[l] = 12.0 * (F[E3(i + 1, j + 1, l)] - 2.0 * F[E3(i + 1, j + 1, l + 1)] + F[E3(i + 1, j + 1, l + 2)]) / hz / hz;
                v[0] -= fzz[E3(i + 1, j + 1, 0)];
                v[Nz - 1] -= fzz[E3(i + 1, j + 1, Nz + 1)];
                csolve(Nz, v, 1, Uz);
                for (int l = 0; l < Nz; l++) fzz[E3(i + 1, j + 1, l + 1)] = v[l];
            }
            free(v);
        }
        ocplx *kp = (ocplx *)malloc(sizeof(ocplx) * Nz), *kpp = (ocplx *)malloc(sizeof(ocplx) * Nz);
        k2_derivs(Nz, hz, &P, k2, kp, kpp);
        double hx2 = hx * hx, hy2 = hy * hy, hz2 = hz * hz;
        /* 2nd_order_partials.c:25-56 and the combination rhs.c:150-166 */
#pragma omp parallel for num_threads(threads)
        for (int l = 0; l < Nz; l++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) {
                    long c0 = E3(i + 1, j + 1, l + 1);
                    ocplx fxxxx = (fxx[c0 - 1] - 2.0 * fxx[c0] + fxx[c0 + 1]) / hx / hx;
                    ocplx fyyyy = (fyy[c0 - X] - 2.0 * fyy[c0] + fyy[c0 + X]) / hy / hy;
                    ocplx fzzzz = (fzz[c0 - XY] - 2.0 * fzz[c0] + fzz[c0 + XY]) / hz / hz;
                    ocplx fxxyy = (fxx[c0 - X] - 2.0 * fxx[c0] + fxx[c0 + X]) / hy / hy;
                    ocplx fxxzz = (fxx[c0 - XY] - 2.0 * fxx[c0] + fxx[c0 + XY]) / hz / hz;
                    ocplx fyyzz = (fyy[c0 - XY] - 2.0 * fyy[c0] + fyy[c0 + XY]) / hz / hz;
                    ocplx first = (1.0 - k2[l + 1] * hz * hz / 20.0) * F[c0];
                    ocplx second = (hz * hz / 12.0) * (fxx[c0] + fyy[c0] + fzz[c0]);
                    ocplx third = (hz * hz * hz * hz / 360.0) * (fxxxx + fyyyy + fzzzz);
                    ocplx fourth = (hz * hz * hz * hz / 90.0) * (fxxyy + fxxzz + fyyzz);
                    ocplx fifth = (hz * hz * hz / 120.0) * kp[l] * (F[c0 + XY] - F[c0 - XY]);
                    rhs[I3(i, j, l)] = first + second + third + fourth + fifth;
                }
        (void)hx2; (void)hy2; (void)hz2;
        free(kp); free(kpp); free(fxx); free(fyy); free(fzz); free(Ux); free(Uy); free(Uz);
    } else {
#pragma omp parallel for num_threads(threads)
        for (int l = 0; l < Nz; l++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) {
                    long c0 = E3(i + 1, j + 1, l + 1);
                    ocplx v = F[c0];
                    if (order == OR_FOURTH)                                   /* rhs.c:34-53 */
                        v = v + (1.0 / 12.0) * (F[c0 - 1] + F[c0 + 1] + F[c0 - X] + F[c0 + X] +
                                                F[c0 - XY] + F[c0 + XY] - 6 * F[c0]);
                    rhs[I3(i, j, l)] = v;
                }
    }
    /* z-face Dirichlet closure: rhs.c:24-29 (2nd), :55-62 (4th), :167-174 (6th) */
    long top = (long)(Nz + 1) * XY;
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            ocplx lo, hi;
            if (order == OR_SECOND) {
                lo = UE[E3(i + 1, j + 1, 0)];
                hi = UE[E3(i + 1, j + 1, Nz + 1)];
            } else {
                lo = am[0] * (UE[E3(i, j, 0)] + UE[E3(i + 2, j, 0)] + UE[E3(i, j + 2, 0)] + UE[E3(i + 2, j + 2, 0)]) +
                     bm[0] * (UE[E3(i, j + 1, 0)] + UE[E3(i + 2, j + 1, 0)]) +
                     cm[0] * (UE[E3(i + 1, j, 0)] + UE[E3(i + 1, j + 2, 0)]) + dm[0] * UE[E3(i + 1, j + 1, 0)];
                hi = ap[Nz - 1] * (UE[top + E3(i, j, 0)] + UE[top + E3(i + 2, j, 0)] + UE[top + E3(i, j + 2, 0)] +
                                   UE[top + E3(i + 2, j + 2, 0)]) +
                     bp[Nz - 1] * (UE[top + E3(i, j + 1, 0)] + UE[top + E3(i + 2, j + 1, 0)]) +
                     cp[Nz - 1] * (UE[top + E3(i + 1, j, 0)] + UE[top + E3(i + 1, j + 2, 0)]) +
                     dp[Nz - 1] * UE[top + E3(i + 1, j + 1, 0)];
            }
            rhs[I3(i, j, 0)] -= lo;
            rhs[I3(i, j, Nz - 1)] -= hi;
        }
    free(UE); free(F); free(k2);
    return 0;
}

/* LU.c:4-81: factors at l + i*Nz + j*Nx*Nz; U holds the reciprocal pivot */
void or3d_LU(int Nx, int Ny, int Nz, const ocplx *coef, ocplx *L, ocplx *U, ocplx *Up) {
    const ocplx *a = coef, *b = coef + Nz, *c = coef + 2 * Nz, *d = coef + 3 * Nz;
    const ocplx *ap = coef + 4 * Nz, *bp = coef + 5 * Nz, *cp = coef + 6 * Nz, *dp = coef + 7 * Nz;
    const ocplx *am = coef + 8 * Nz, *bm = coef + 9 * Nz, *cm = coef + 10 * Nz, *dm = coef + 11 * Nz;
    long Nxz = (long)Nx * Nz;
#pragma omp parallel for
    for (int j = 0; j < Ny; j++) {
        double ly = 2.0 * cos((j + 1) * M_PI / (Ny + 1));
        for (int i = 0; i < Nx; i++) {
            double lx = 2.0 * cos((i + 1) * M_PI / (Nx + 1));
            double lxy = lx * ly;
            long m = (long)i * Nz + (long)j * Nxz;
            L[m] = 0.0;
            U[m] = 1.0 / (a[0] * lxy + b[0] * lx + c[0] * ly + d[0]);
            Up[m] = ap[0] * lxy + bp[0] * lx + cp[0] * ly + dp[0];
            for (int l = 1; l < Nz; l++) {
                ocplx e3 = a[l] * lxy + b[l] * lx + c[l] * ly + d[l];
                ocplx e3m = am[l] * lxy + bm[l] * lx + cm[l] * ly + dm[l];
                L[l + m] = e3m * U[l - 1 + m];
                U[l + m] = 1.0 / (e3 - Up[l - 1 + m] * L[l + m]);
                Up[l + m] = ap[l] * lxy + bp[l] * lx + cp[l] * ly + dp[l];
            }
        }
    }
}

/* 2D DST of one z-layer (DST.c:148-204): x lines then y lines, orthonormal */
static void dst2d_layer(const or_dstplan *px, const or_dstplan *py, int Nx, int Ny,
                        const ocplx *in, long istride, ocplx *tmp, ocplx *out) {
    or_dst_lines_plan(px, Ny, in, (long)Nx * istride, istride, tmp, Nx, 1, 1);
    or_dst_lines_plan(py, Nx, tmp, 1, Nx, out, 1, Nx, 1);
}

/* solver.c:263-413 (CPU branch), plans created inside the call (:289-331) */
void or3d_solver(int Nx, int Ny, int Nz, const ocplx *rhs, const ocplx *L,
                 const ocplx *U, const ocplx *Up, ocplx *sol, int threads) {
    or_dstplan *px = or_dst_plan(Nx), *py = or_dst_plan(Ny);        /* :289-331 */
    or3d_solver_planned(px, py, Nx, Ny, Nz, rhs, L, U, Up, sol, threads);
    or_dst_plan_free(px);
    or_dst_plan_free(py);
}

/* the same solve with caller-held plans (steady state) */
void or3d_solver_planned(const or_dstplan *px, const or_dstplan *py, int Nx, int Ny, int Nz, const ocplx *rhs,
                         const ocplx *L, const ocplx *U, const ocplx *Up, ocplx *sol, int threads) {
    if (threads < 1) threads = 1;
    long Nxy = (long)Nx * Ny, Nxz = (long)Nx * Nz;
    ocplx *rhat = (ocplx *)malloc(sizeof(ocplx) * Nxy * Nz);
    ocplx *xhat = (ocplx *)malloc(sizeof(ocplx) * Nxy * Nz);
#pragma omp parallel num_threads(threads)
    {
        ocplx *tmp = (ocplx *)malloc(sizeof(ocplx) * Nxy);
#pragma omp for
        for (int l = 0; l < Nz; l++) dst2d_layer(px, py, Nx, Ny, rhs + l * Nxy, 1, tmp, rhat + l * Nxy);   /* :338-350 */
#pragma omp for
        for (int j = 0; j < Ny; j++) {                                                         /* :356-369 */
            ocplx *ya = (ocplx *)malloc(sizeof(ocplx) * Nz);
            for (int i = 0; i < Nx; i++) {
                long m = (long)i * Nz + (long)j * Nxz;
                ya[0] = rhat[i + (long)j * Nx];
                for (int l = 1; l < Nz; l++) ya[l] = rhat[i + (long)j * Nx + l * Nxy] - L[l + m] * ya[l - 1];
                xhat[Nz - 1 + m] = ya[Nz - 1] * U[Nz - 1 + m];
                for (int l = Nz - 2; l >= 0; l--) xhat[l + m] = (ya[l] - Up[l + m] * xhat[l + 1 + m]) * U[l + m];
            }
            free(ya);
        }
#pragma omp for
        for (int l = 0; l < Nz; l++) dst2d_layer(px, py, Nx, Ny, xhat + l, Nz, tmp, sol + l * Nxy);       /* :376-388 */
        free(tmp);
    }
    free(rhat);
    free(xhat);
}

/* residual.c:42-96 */
void or3d_residual(int Nx, int Ny, int Nz, const ocplx *coef, const ocplx *sol,
                   const ocplx *rhs, ocplx *res) {
    const ocplx *a = coef, *b = coef + Nz, *c = coef + 2 * Nz, *d = coef + 3 * Nz;
    const ocplx *ap = coef + 4 * Nz, *bp = coef + 5 * Nz, *cp = coef + 6 * Nz, *dp = coef + 7 * Nz;
    const ocplx *am = coef + 8 * Nz, *bm = coef + 9 * Nz, *cm = coef + 10 * Nz, *dm = coef + 11 * Nz;
    long Nxy = (long)Nx * Ny, X = Nx + 2, XY = (long)(Nx + 2) * (Ny + 2);
    ocplx *s = (ocplx *)calloc(XY * (Nz + 2), sizeof(ocplx));
    for (int l = 0; l < Nz; l++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) s[E3(i + 1, j + 1, l + 1)] = sol[I3(i, j, l)];
#pragma omp parallel for
    for (int l = 0; l < Nz; l++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                ocplx v[3];
                const ocplx *cA[3] = {am, a, ap}, *cB[3] = {bm, b, bp}, *cC[3] = {cm, c, cp}, *cD[3] = {dm, d, dp};
                for (int q = 0; q < 3; q++) {
                    long c0 = E3(i + 1, j + 1, l + q);
                    v[q] = cD[q][l] * s[c0] + cA[q][l] * (s[c0 + 1 + X] + s[c0 - 1 + X] + s[c0 + 1 - X] + s[c0 - 1 - X]) +
                           cB[q][l] * (s[c0 + 1] + s[c0 - 1]) + cC[q][l] * (s[c0 + X] + s[c0 - X]);
                }
                res[I3(i, j, l)] = (v[2] + v[1] + v[0]) - rhs[I3(i, j, l)];
            }
    free(s);
}

/* Inputs of the GPU right-hand-side assembly for the same problem:
 * f_e   : hz^2 f on the extended (Nx+2)(Ny+2)(Nz+2) grid
 * u_e   : u on the extended grid (Dirichlet data; only z faces are used)
 * fxx_f : hz^2 f_xx on the x faces, [face][l][j] (2*(Ny+2)*(Nz+2))
 * fyy_f : hz^2 f_yy on the y faces, [face][l][i] (2*(Nx+2)*(Nz+2))
 * fzz_f : hz^2 f_zz on the z faces, [face][j][i] (2*(Nx+2)*(Ny+2))
 * k2_e  : k^2 at z0 + l*hz, l = 0..Nz+1
 * k2z_b : (k^2)_z at z0, z1 and (k^2)_zz at z0, z1 */
void or3d_fields(int Nx, int Ny, int Nz, int kind, const double *p,
                 ocplx *f_e, ocplx *u_e, ocplx *fxx_f, ocplx *fyy_f,
                 ocplx *fzz_f, ocplx *k2_e, ocplx *k2z_b) {
    pr3 P;
    pr3_init(&P, kind, p);
    long X = Nx + 2, XY = (long)(Nx + 2) * (Ny + 2);
    double hx = (P.x1 - P.x0) / (Nx + 1), hy = (P.y1 - P.y0) / (Ny + 1), hz = (P.z1 - P.z0) / (Nz + 1);
    double h2 = hz * hz;
#pragma omp parallel for
    for (int l = 0; l < Nz + 2; l++)
        for (int j = 0; j < Ny + 2; j++)
            for (int i = 0; i < Nx + 2; i++) {
                double x = P.x0 + i * hx, y = P.y0 + j * hy, z = P.z0 + l * hz;
                f_e[E3(i, j, l)] = h2 * pr3_f(&P, x, y, z);
                u_e[E3(i, j, l)] = pr3_u(&P, x, y, z);
            }
    long Fx = (long)(Ny + 2) * (Nz + 2), Fy = (long)(Nx + 2) * (Nz + 2), Fz = XY;
    for (int l = 0; l < Nz + 2; l++)
        for (int j = 0; j < Ny + 2; j++) {
            double y = P.y0 + j * hy, z = P.z0 + l * hz;
            fxx_f[j + (long)l * (Ny + 2)] = h2 * pr3_f_xx(&P, P.x0, y, z);
            fxx_f[Fx + j + (long)l * (Ny + 2)] = h2 * pr3_f_xx(&P, P.x1, y, z);
        }
    for (int l = 0; l < Nz + 2; l++)
        for (int i = 0; i < Nx + 2; i++) {
            double x = P.x0 + i * hx, z = P.z0 + l * hz;
            fyy_f[i + (long)l * (Nx + 2)] = h2 * pr3_f_yy(&P, x, P.y0, z);
            fyy_f[Fy + i + (long)l * (Nx + 2)] = h2 * pr3_f_yy(&P, x, P.y1, z);
        }
    for (int j = 0; j < Ny + 2; j++)
        for (int i = 0; i < Nx + 2; i++) {
            double x = P.x0 + i * hx, y = P.y0 + j * hy;
            fzz_f[i + (long)j * (Nx + 2)] = h2 * pr3_f_zz(&P, x, y, P.z0);
            fzz_f[Fz + i + (long)j * (Nx + 2)] = h2 * pr3_f_zz(&P, x, y, P.z1);
        }
    for (int l = 0; l < Nz + 2; l++) k2_e[l] = pr3_k2(&P, P.z0 + l * hz);
    k2z_b[0] = pr3_k2_z(&P, P.z0);
    k2z_b[1] = pr3_k2_z(&P, P.z1);
    k2z_b[2] = pr3_k2_zz(&P, P.z0);
    k2z_b[3] = pr3_k2_zz(&P, P.z1);
}
