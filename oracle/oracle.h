/*
 * CPU FP64 oracle for the compact-scheme Helmholtz direct solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle that the GPU
 * library is compared against.  Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load it; the product never does.
 *
 * It restates, in plain C99 + OpenMP, the reference host pipeline of
 * gonzrona/hackathon:
 *   2D  2D_Program/main_functions/{coefficients.c,rhs.c,LU.c,residual.c},
 *       2D_Program/partial_derivatives/ (all files),
 *       2D_Program/solver_functions/{solver.c,DST.c}
 *   3D  3D_Program/main_functions/{coefficients.c,rhs.c,LU.c,residual.c},
 *       3D_Program/partial_derivatives/ (all files),
 *       3D_Program/solver_functions/solver.c:263-413, DST.c:148-204
 * with the same array layouts (row-major i + j*Nx (+ l*Nx*Ny), interleaved
 * complex double) and the same arithmetic order where it matters.
 *
 * Differences from the reference, on purpose:
 *  - FFTW is not available, so the DST-I is computed with the oracle's own
 *    complex FFT (radix-2 for power-of-two lengths, Bluestein otherwise) of
 *    the odd extension of each complex row.  This equals the reference's two
 *    real r2c transforms of the zero-padded Re/Im rows up to roundoff.
 *  - The problem (k profile, manufactured solution) is chosen at run time
 *    (problem kind + parameters) instead of being compiled in user_input.c.
 *  - Dirichlet data may be complex (the 2D reference only allows real data).
 *  - The reference's uninitialised 3D ghost cells are not reproduced: all
 *    ghosts of the compact second derivatives are set from the problem's
 *    analytic boundary values.
 *
 * Parity status of every function is listed in the test suite; functions
 * without an independent check would be marked "parity unpinned" here.
 * None currently is.
 */
#ifndef HH_ORACLE_H
#define HH_ORACLE_H

#include <complex.h>
#include <stddef.h>

typedef double _Complex ocplx;

enum { OR_SECOND = 2, OR_FOURTH = 4, OR_SIXTH = 6 };

/* Problem kinds.
 * 2D kind 0 (REF2D): 2D_Program/user_input.c:4-58, k(y)=y^2+iy,
 *    u=sin x sin y on [0.5, pi-0.5]^2.  params unused.
 * 2D kind 1 (SINE2D): constant k = p[0] + i p[1]; u = e^{i p[4]} sin(p[2] pi x)
 *    sin(p[3] pi y) on the unit square.
 * 3D kind 0 (REF3D): 3D_Program/user_input.c:5-44, k(z)=A-B sin(Cz),
 *    u = sin(gamma x) sin(beta y) exp(-k(z)/C) on [0,pi]^3,
 *    p = {A, B, C, gamma}; beta = sqrt(A^2+B^2-gamma^2).
 * 3D kind 1 (SINE3D): constant k = p[0] + i p[1]; u = (p[5] + i p[6])
 *    sin(p[2] pi x) sin(p[3] pi y) sin(p[4] pi z) on the unit cube.
 */
enum { OR_REF = 0, OR_SINE = 1 };

#ifdef __cplusplus
extern "C" {
#endif

/* ---- DST-I (orthonormal), 2D_Program/solver_functions/DST.c:6-44 ---- */
/* out[r*N + k] = sqrt(2/(N+1)) * sum_n in[r*N + n] sin(pi (n+1)(k+1)/(N+1)) */
void or_dst1_rows(int N, long nrows, const ocplx *in, ocplx *out, int threads);

/* ---- 2D ---- */
/* Domain of a 2D problem: x0,x1,y0,y1 */
void or2d_domain(int kind, const double *p, double dom[4]);
/* coefficients.c + rhs.c + partial derivatives: stencil coefficient vectors
 * (length Ny), the compact right-hand side (Nx*Ny) and the analytic
 * solution on the interior (Nx*Ny). */
int or2d_setup(int order, int Nx, int Ny, int kind, const double *p,
               ocplx *bp, ocplx *ap, ocplx *b, ocplx *a, ocplx *bm, ocplx *am,
               ocplx *rhs, ocplx *sol_analytic, int threads);
/* LU.c:3-46: mode-major factors L,U,Up (index j + i*Ny). */
void or2d_LU(int Nx, int Ny, const ocplx *bp, const ocplx *ap, const ocplx *b,
             const ocplx *a, const ocplx *bm, const ocplx *am,
             ocplx *L, ocplx *U, ocplx *Up, int threads);
/* solver.c:14-117: sol = S_x T^{-1} S_x rhs, using the stored factors. */
void or2d_solver(int Nx, int Ny, const ocplx *rhs, const ocplx *L,
                 const ocplx *U, const ocplx *Up, ocplx *sol, int threads);
/* Steady state: the same solves with caller-held DST plans (plan creation
 * outside the timed call).  or_plan(N) makes the DST-I plan of length N. */
typedef struct or_dstplan or_dstplan;
or_dstplan *or_dst_plan(int N);
void or_dst_plan_free(or_dstplan *plan);
void or2d_solver_planned(const or_dstplan *plan, int Nx, int Ny, const ocplx *rhs, const ocplx *L,
                         const ocplx *U, const ocplx *Up, ocplx *sol, int threads);
/* residual.c:3-34: res = A sol - rhs (9-point operator, zero ghosts). */
void or2d_residual(int Nx, int Ny, const ocplx *bp, const ocplx *ap,
                   const ocplx *b, const ocplx *a, const ocplx *bm,
                   const ocplx *am, const ocplx *sol, const ocplx *rhs,
                   ocplx *res);
/* Right-hand-side ingredients used by the product's GPU assembly kernels:
 * f on the extended (Nx+2)x(Ny+2) grid, and analytic boundary data. */
void or2d_fields(int Nx, int Ny, int kind, const double *p, ocplx *f_e,
                 ocplx *u_e, ocplx *fy_b, ocplx *fxx_b, ocplx *fyy_b,
                 ocplx *k2_e, ocplx *k2y_b);

/* ---- 3D ---- */
void or3d_domain(int kind, const double *p, double dom[6]);
int or3d_setup(int order, int Nx, int Ny, int Nz, int kind, const double *p,
               ocplx *coef /* 12*Nz: a,b,c,d,ap,bp,cp,dp,am,bm,cm,dm */,
               ocplx *rhs, ocplx *sol_analytic, int threads);
/* LU.c:4-81: factors indexed l + i*Nz + j*Nx*Nz, U is the reciprocal pivot */
void or3d_LU(int Nx, int Ny, int Nz, const ocplx *coef, ocplx *L, ocplx *U,
             ocplx *Up);
/* solver.c:263-413 (CPU branch) */
void or3d_solver(int Nx, int Ny, int Nz, const ocplx *rhs, const ocplx *L,
                 const ocplx *U, const ocplx *Up, ocplx *sol, int threads);
void or3d_solver_planned(const or_dstplan *px, const or_dstplan *py, int Nx, int Ny, int Nz, const ocplx *rhs,
                         const ocplx *L, const ocplx *U, const ocplx *Up, ocplx *sol, int threads);
/* residual.c:42-96 (27-point operator, zero ghosts) */
void or3d_residual(int Nx, int Ny, int Nz, const ocplx *coef, const ocplx *sol,
                   const ocplx *rhs, ocplx *res);
/* extended fields for the GPU RHS assembly: f (times hz^2) on the extended
 * grid, u on the extended grid, the three face arrays of analytic second
 * derivatives (times hz^2) and k^2 on the extended z grid. */
void or3d_fields(int Nx, int Ny, int Nz, int kind, const double *p,
                 ocplx *f_e, ocplx *u_e, ocplx *fxx_f, ocplx *fyy_f,
                 ocplx *fzz_f, ocplx *k2_e, ocplx *k2z_b);

/* ---- norms, 2D_Program/helpers_functions/norms.c:3-13,
 *      3D_Program/helpers_functions/norms.c:15-24 ---- */
double or_normInf(long n, const ocplx *x);
double or_normL2(long n, const ocplx *x);

/* ---- wall clock around a full oracle solve (CPU baseline) ---- */
double or_wtime(void);

#ifdef __cplusplus
}
#endif
#endif
