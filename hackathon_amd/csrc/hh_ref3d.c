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
