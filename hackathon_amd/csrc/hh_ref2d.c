/*
 * Drop-in solver(System), forwardDST and reverseDST for the 2D program
 * (include/hh_ref2d.h), in plain C over the C ABI of hh_solver.h.
 *
 * Replaces 2D_Program/solver_functions/solver.c:14-117 and DST.c:6-44.
 * The reference builds FFTW plans and scratch on every call; here one
 * context and one plan persist across calls and the plan is rebuilt only
 * when the grid or the stencil vectors change.
 */
#include "hh_ref2d.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hh_solver.h"

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static hh_context *g_ctx;
static hh_plan *g_plan;
static int g_Nx, g_Ny;
static double _Complex *g_coef; /* copy of bp, ap, b, a, bm, am the plan was built from */
static __thread int g_status;

int hh_ref2d_last_status(void) { return g_status; }

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
    int Nx = s->lat.Nx, Ny = s->lat.Ny;
    const double _Complex *v[6] = {s->bp, s->ap, s->b, s->a, s->bm, s->am};
    if (g_plan && g_Nx == Nx && g_Ny == Ny) {
        int same = 1;
        for (int q = 0; q < 6 && same; q++) same = !memcmp(g_coef + (size_t)q * Ny, v[q], sizeof(double _Complex) * Ny);
        if (same) return HH_OK;
    }
    if (g_plan) { hh_plan_destroy(g_plan); g_plan = NULL; }
    free(g_coef);
    g_coef = (double _Complex *)malloc(sizeof(double _Complex) * 6 * (size_t)Ny);
    if (!g_coef) return HH_ERR_NOMEM;
    for (int q = 0; q < 6; q++) memcpy(g_coef + (size_t)q * Ny, v[q], sizeof(double _Complex) * Ny);
    g_Nx = Nx;
    g_Ny = Ny;
    return hh_plan2d_create(g_ctx, Nx, Ny, (const double *)s->bp, (const double *)s->ap, (const double *)s->b,
                            (const double *)s->a, (const double *)s->bm, (const double *)s->am, HH_PREC_F64, &g_plan);
}

void solver(System sys) {
    pthread_mutex_lock(&g_mu);
    int rc = context();
    if (!rc) rc = plan_for(&sys);
    if (!rc) rc = hh_solve_host(g_plan, (const double *)sys.rhs, (double *)sys.sol, 1);
    pthread_mutex_unlock(&g_mu);
    report("solver", rc);
}

void forwardDST(System sys, DSTN dst, double _Complex *rhs, double _Complex *rhat, fftw_plan plan, double *in,
                fftw_complex *out, fftw_plan plan2, double *in2, fftw_complex *out2) {
    (void)dst; (void)plan; (void)in; (void)out; (void)plan2; (void)in2; (void)out2;
    pthread_mutex_lock(&g_mu);
    int rc = context();
    if (!rc) rc = hh_dst_rows_host(g_ctx, sys.lat.Nx, sys.lat.Ny, (const double *)rhs, (double *)rhat);
    pthread_mutex_unlock(&g_mu);
    report("forwardDST", rc);
}

void reverseDST(System sys, DSTN dst, double _Complex *xhat, double _Complex *sol, fftw_plan plan, double *in,
                fftw_complex *out, fftw_plan plan2, double *in2, fftw_complex *out2) {
    (void)dst; (void)plan; (void)in; (void)out; (void)plan2; (void)in2; (void)out2;
    int Nx = sys.lat.Nx, Ny = sys.lat.Ny;
    /* the reference gathers xhat[j + i*Ny] for row j (mode-major input) */
    double _Complex *t = (double _Complex *)malloc(sizeof(double _Complex) * (size_t)Nx * Ny);
    if (!t) { report("reverseDST", HH_ERR_NOMEM); return; }
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++) t[i + (size_t)j * Nx] = xhat[j + (size_t)i * Ny];
    pthread_mutex_lock(&g_mu);
    int rc = context();
    if (!rc) rc = hh_dst_rows_host(g_ctx, Nx, Ny, (const double *)t, (double *)sol);
    pthread_mutex_unlock(&g_mu);
    free(t);
    report("reverseDST", rc);
}
