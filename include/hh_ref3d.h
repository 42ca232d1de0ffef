/*
 * hh_ref3d.h -- drop-in replacement of the 3D program's solver functions.
 *
 * Link a program built from 3D_Program/main_functions, partial_derivatives,
 * helpers_functions and user_input.c against libhh_ref3d.so instead of
 * compiling 3D_Program/solver_functions/{solver.c,DST.c,cuda_kernels.cu}.
 *
 * Types mirror 3D_Program/headers/structs.h:14-33 member for member.  The
 * FFTW types only keep the DST() signature of 3D_Program/solver_functions/
 * DST.c:148 (FFTW branch); the GPU path ignores plans and buffers.  Do not
 * include together with hh_ref2d.h.
 */
#ifndef HH_REF3D_H
#define HH_REF3D_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

enum Order { second, fourth, sixth };

typedef struct {
    int Nx, Ny, Nz, Nxyz;
    double hx, hy, hz, x0, x1, y0, y1, z0, z1;
} Lattice;

typedef struct {
    Lattice lat;
    enum Order order;
    double _Complex *a;
    double _Complex *b, *c, *d, *ap, *bp, *cp, *dp, *am, *bm, *cm, *dm;
    double _Complex *k_bg_ext, *k2_bg_ext;
    double A, B, C, beta, gamma;
    double _Complex *sol_analytic, *rhs;
    double _Complex *L, *U, *Up;
    double _Complex *sol, *res, *error;
} System;

typedef struct hh_fftw_plan_s *fftw_plan;
typedef double fftw_complex[2];

/* 3D_Program/headers/prototypes.h:14 -- sys.sol = A^{-1} sys.rhs from the
 * twelve stencil vectors sys.a..sys.dm (length Nz); L/U/Up are not read. */
void solver(System sys);

/* 3D_Program/solver_functions/DST.c:148 -- bhat = 2D DST-I (x then y,
 * scaled 2/sqrt((Nx+1)(Ny+1))) of the real Nx*Ny layer b_2D.  Unlike the
 * reference, b_2D is left unchanged. */
void DST(int Nx, int Ny, double *b_2D, double *bhat, fftw_plan p1, double *in1, fftw_complex *out1, fftw_plan p2,
         double *in2, fftw_complex *out2);

/* 3D_Program/main_functions/rhs.c:198-294 with the partial derivatives:
 * the reference problem u = sin(gamma x) sin(beta y) exp(-k(z)/C),
 * k(z) = A - B sin(Cz) (parameters sys.A, B, C, gamma, beta and the
 * background sys.k_bg_ext, sys.k2_bg_ext), sampled on the host as the
 * reference's user-input blocks do, assembled on the GPU; fills
 * sys.sol_analytic and sys.rhs.  Needs the stencil vectors. */
void rhs(System sys);

/* 3D_Program/main_functions/residual.c:42-96: sys.res = A sys.sol - sys.rhs
 * (27-point operator), sys.error = sys.sol - sys.sol_analytic. */
void residual(System sys);

/* Status of the last call on this thread (0 = ok); no reference counterpart. */
int hh_ref3d_last_status(void);

#ifdef __cplusplus
}
#endif
#endif
