/*
 * hh_ref2d.h -- drop-in replacement of the 2D program's solver functions.
 *
 * Link a program built from 2D_Program/main_functions, partial_derivatives,
 * helpers_functions and user_input.c against libhh_ref2d.so instead of
 * compiling 2D_Program/solver_functions/{solver.c,DST.c}: the symbols and
 * the System layout below are those of the reference, the work runs on the
 * MI355X through the C ABI of hh_solver.h.
 *
 * Types mirror 2D_Program/headers/structs.h:12-30 member for member (the
 * layout is the interface).  fftw_plan / fftw_complex only keep the
 * forwardDST / reverseDST signatures of 2D_Program/solver_functions/
 * solver.c:10-11; the GPU path ignores the plan and buffer arguments.
 * Do not include together with hh_ref3d.h (the two programs' System types
 * differ, as in the reference).
 */
#ifndef HH_REF2D_H
#define HH_REF2D_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

enum Order { second, fourth, sixth };

typedef struct {
    int Nx, N;
    double coef;
} DSTN;

typedef struct {
    int Nx, Ny, Nxy;
    double hx, hy, x0, x1, y0, y1;
} Lattice;

typedef struct {
    Lattice lat;
    int threadCount;
    enum Order order;
    double _Complex *k, *rhs, *L, *U, *Up, *sol, *sol_analytic, *res, *error;
    double _Complex *bp, *ap, *b, *a, *bm, *am;
} System;

typedef struct hh_fftw_plan_s *fftw_plan;
typedef double fftw_complex[2];

/* 2D_Program/headers/prototypes.h:25 -- sys.sol = A^{-1} sys.rhs.
 * Reads sys.lat.Nx/Ny, sys.rhs and the stencil vectors sys.bp..sys.am
 * (length Ny); sys.L/U/Up are not read (the factors are rebuilt on the GPU).
 * Errors are printed to stderr, as the reference's CUDA_RT_CALL does, and
 * kept for hh_ref2d_last_status(). */
void solver(System sys);

/* 2D_Program/solver_functions/DST.c:6 -- rhat = DST-I of every row of rhs
 * (Ny rows of length Nx, orthonormal, sqrt(2/(Nx+1)) scaling). */
void forwardDST(System sys, DSTN dst, double _Complex *rhs, double _Complex *rhat, fftw_plan plan, double *in,
                fftw_complex *out, fftw_plan plan2, double *in2, fftw_complex *out2);

/* 2D_Program/solver_functions/DST.c:26 -- row j of sol = DST-I of the
 * mode-major column xhat[j + i*Ny], i = 0..Nx-1. */
void reverseDST(System sys, DSTN dst, double _Complex *xhat, double _Complex *sol, fftw_plan plan, double *in,
                fftw_complex *out, fftw_plan plan2, double *in2, fftw_complex *out2);

/* 2D_Program/headers/prototypes.h:23, main_functions/rhs.c:100-150 with the
 * partial derivatives: fills sys.k, sys.sol_analytic and the compact
 * right-hand side sys.rhs (Dirichlet closure included) for sys.order.
 * Samples the program's own user-input functions k, k2_y, f, f_y, f_xx,
 * f_yy and analytic_solution (2D_Program/user_input.c) on the host and
 * assembles on the GPU.  Needs the stencil vectors from coefficients(). */
void rhs(System sys);

/* 2D_Program/main_functions/residual.c:3-34: sys.res = A sys.sol - sys.rhs
 * (9-point operator, zero ghosts), sys.error = sys.sol - sys.sol_analytic. */
void residual(System sys);

/* Status of the last call on this thread (HH_OK = 0); no reference
 * counterpart, the reference's functions return void. */
int hh_ref2d_last_status(void);

#ifdef __cplusplus
}
#endif
#endif
