/*
 * Drop-in solver(System) and DST() for the 3D program (include/hh_ref3d.h),
 * in plain C over the C ABI of hh_solver.h.
 *
 * Replaces 3D_Program/solver_functions/solver.c:55-249 (CUDA) and :263-413
 * (FFTW), DST.c:13-204 and cuda_kernels.cu.  One context and one plan
 * persist across calls; the plan is rebuilt when the grid or the twelve
 * stencil vectors change.
 */
#include "hh_ref3d.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hh_solver.h"

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static hh_context *g_ctx;
static hh_plan *g_plan;
static int g_N[3];
static double _Complex *g_coef;
static __thread int g_status;

int hh_ref3d_last_status(void) { return g_status; }

static int report(const char *fn, int rc) {
    g_status = rc;
    if (rc != HH_OK) fprintf(stderr, "ERROR: %s failed: %s (%s)\n", fn, hh_status_string(rc), hh_last_error());
    return rc;
}

static int context(void) {
    if (g_ctx) return HH_OK;
    const char *d = getenv("HH_DEVICE");
    return hh_context_create(d ? atoi(d) : 0, &g_ctx);
}

static int plan_for(const System *s) {
    int Nx = s->lat.Nx, Ny = s->lat.Ny, Nz = s->lat.Nz;
    const double _Complex *v[12] = {s->a, s->b, s->c, s->d, s->ap, s->bp, s->cp, s->dp, s->am, s->bm, s->cm, s->dm};
    if (g_plan && g_N[0] == Nx && g_N[1] == Ny && g_N[2] == Nz) {
        int same = 1;
        for (int q = 0; q < 12 && same; q++)
            same = !memcmp(g_coef + (size_t)q * Nz, v[q], sizeof(double _Complex) * Nz);
        if (same) return HH_OK;
    }
    if (g_plan) { hh_plan_destroy(g_plan); g_plan = NULL; }
    free(g_coef);
    g_coef = (double _Complex *)malloc(sizeof(double _Complex) * 12 * (size_t)Nz);
    if (!g_coef) return HH_ERR_NOMEM;
    const double *ptr[12];
    for (int q = 0; q < 12; q++) {
        memcpy(g_coef + (size_t)q * Nz, v[q], sizeof(double _Complex) * Nz);
        ptr[q] = (const double *)v[q];
    }
    g_N[0] = Nx; g_N[1] = Ny; g_N[2] = Nz;
    return hh_plan3d_create(g_ctx, Nx, Ny, Nz, ptr, HH_PREC_F64, &g_plan);
}

void solver(System sys) {
    pthread_mutex_lock(&g_mu);
    int rc = context();
    if (!rc) rc = plan_for(&sys);
    if (!rc) rc = hh_solve_host(g_plan, (const double *)sys.rhs, (double *)sys.sol, 1);
    pthread_mutex_unlock(&g_mu);
    report("solver", rc);
}

void DST(int Nx, int Ny, double *b_2D, double *bhat, fftw_plan p1, double *in1, fftw_complex *out1, fftw_plan p2,
         double *in2, fftw_complex *out2) {
    (void)p1; (void)in1; (void)out1; (void)p2; (void)in2; (void)out2;
    size_t n = (size_t)Nx * Ny;
    double *z = (double *)calloc(2 * n, sizeof(double));
    if (!z) { report("DST", HH_ERR_NOMEM); return; }
    for (size_t q = 0; q < n; q++) z[2 * q] = b_2D[q];
    pthread_mutex_lock(&g_mu);
    int rc = context();
    if (!rc) rc = hh_dst2d_layers_host(g_ctx, Nx, Ny, 1, z, z);
    pthread_mutex_unlock(&g_mu);
    if (!rc)
        for (size_t q = 0; q < n; q++) bhat[q] = z[2 * q];
    free(z);
    report("DST", rc);
}

static int order_of(enum Order o) { return o == sixth ? 6 : o == fourth ? 4 : 2; }

void rhs(System sys) {
    const int Nx = sys.lat.Nx, Ny = sys.lat.Ny, Nz = sys.lat.Nz;
    const long X = Nx + 2, Y = Ny + 2, Z = Nz + 2, XY = X * Y;
    const double hx = sys.lat.hx, hy = sys.lat.hy, hz = sys.lat.hz, z0 = sys.lat.z0, z1 = sys.lat.z1;
    const double A = sys.A, B = sys.B, C = sys.C, gm = sys.gamma, bt = sys.beta;
    double _Complex *ue = malloc(sizeof(double _Complex) * XY * Z), *fe = malloc(sizeof(double _Complex) * XY * Z);
    double _Complex *fxx = malloc(sizeof(double _Complex) * 2 * Y * Z), *fyy = malloc(sizeof(double _Complex) * 2 * X * Z);
    double _Complex *fzz = malloc(sizeof(double _Complex) * 2 * XY), k2d[4];
    double *fix = malloc(sizeof(double) * X), *fiy = malloc(sizeof(double) * Y);
    if (!ue || !fe || !fxx || !fyy || !fzz || !fix || !fiy) {
        free(ue); free(fe); free(fxx); free(fyy); free(fzz); free(fix); free(fiy);
        report("rhs", HH_ERR_NOMEM);
        return;
    }
    /* user-input block of rhs.c:229-252 */
    for (long i = 0; i < X; i++) fix[i] = sin(gm * (sys.lat.x0 + i * hx));
    for (long j = 0; j < Y; j++) fiy[j] = sin(bt * (sys.lat.y0 + j * hy));
    for (long l = 0; l < Z; l++) {
        double z = z0 + l * hz;
        double _Complex fiz = cexp(-1.0 * sys.k_bg_ext[l] / C);
        for (long j = 0; j < Y; j++)
            for (long i = 0; i < X; i++) {
                long q = i + j * X + l * XY;
                ue[q] = fiz * fix[i] * fiy[j];
                fe[q] = -1.0 * hz * hz * B * (2.0 * A + C) * sin(C * z) * ue[q];
            }
    }
    /* user-input block of rhs.c:90-109 (boundary second derivatives) */
    for (long l = 0; l < Z; l++)
        for (long j = 0; j < Y; j++) {
            fxx[j + l * Y] = -1.0 * gm * gm * fe[j * X + l * XY];
            fxx[Y * Z + j + l * Y] = -1.0 * gm * gm * fe[(X - 1) + j * X + l * XY];
        }
    for (long l = 0; l < Z; l++)
        for (long i = 0; i < X; i++) {
            fyy[i + l * X] = -1.0 * bt * bt * fe[i + l * XY];
            fyy[X * Z + i + l * X] = -1.0 * bt * bt * fe[i + (Y - 1) * X + l * XY];
        }
    double s0 = sin(C * z0), c0 = cos(C * z0), s1 = sin(C * z1), c1 = cos(C * z1);
    double g0 = B * C * s0 * s0 + s0 * (C * C - B * B * c0 * c0) - 2.0 * B * C * c0 * c0;
    double g1 = B * C * s1 * s1 + s1 * (C * C - B * B * c1 * c1) - 2.0 * B * C * c1 * c1;
    for (long j = 0; j < Y; j++)
        for (long i = 0; i < X; i++) {
            fzz[i + j * X] = hz * hz * B * (2.0 * A + C) * ue[i + j * X] * g0;
            fzz[XY + i + j * X] = hz * hz * B * (2.0 * A + C) * ue[i + j * X + (Z - 1) * XY] * g1;
        }
    /* (k^2)_z, (k^2)_zz at the z faces: 4th_order_partials.c:314-315, 359-360 */
    k2d[0] = -2.0 * B * C * c0 * (A - B * s0);
    k2d[1] = -2.0 * B * C * c1 * (A - B * s1);
    k2d[2] = -2.0 * B * C * C * (B * s0 * s0 - A * s0 - B * c0 * c0);
    k2d[3] = -2.0 * B * C * C * (B * s1 * s1 - A * s1 - B * c1 * c1);
    for (long l = 0; l < Nz; l++)                                     /* rhs.c:261-267 */
        for (long j = 0; j < Ny; j++)
            for (long i = 0; i < Nx; i++)
                sys.sol_analytic[i + j * Nx + l * (long)Nx * Ny] = ue[(i + 1) + (j + 1) * X + (l + 1) * XY];
    pthread_mutex_lock(&g_mu);
    int rc = context();
    if (!rc) rc = plan_for(&sys);
    if (!rc)
        rc = hh_rhs3d_host(g_plan, order_of(sys.order), hx, hy, hz, (const double *)fe, (const double *)ue,
                           (const double *)fxx, (const double *)fyy, (const double *)fzz,
                           (const double *)sys.k2_bg_ext, (const double *)k2d, (double *)sys.rhs);
    pthread_mutex_unlock(&g_mu);
    free(ue); free(fe); free(fxx); free(fyy); free(fzz); free(fix); free(fiy);
    report("rhs", rc);
}

void residual(System sys) {
    pthread_mutex_lock(&g_mu);
    int rc = context();
    if (!rc) rc = plan_for(&sys);
    if (!rc) rc = hh_residual_host(g_plan, (const double *)sys.sol, (const double *)sys.rhs, (double *)sys.res, 1);
    pthread_mutex_unlock(&g_mu);
    if (!rc)                                                           /* residual.c:81 */
        for (long i = 0; i < (long)sys.lat.Nxyz; i++) sys.error[i] = sys.sol[i] - sys.sol_analytic[i];
    report("residual", rc);
}
