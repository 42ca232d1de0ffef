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

/* The program's user-input functions (2D_Program/headers/prototypes.h:4-16).
 * Weak, so the library loads in programs that do not define them. */
extern double _Complex k(double y) __attribute__((weak));
extern double _Complex k2_y(double y) __attribute__((weak));
extern double _Complex f(double x, double y) __attribute__((weak));
extern double _Complex f_y(double x, double y) __attribute__((weak));
extern double _Complex f_xx(double x, double y) __attribute__((weak));
extern double _Complex f_yy(double x, double y) __attribute__((weak));
extern double analytic_solution(double x, double y) __attribute__((weak));
/* optional complex Dirichlet data; the reference's analytic_solution is real */
extern double _Complex analytic_solution_c(double x, double y) __attribute__((weak));

static double _Complex u_at(double x, double y) {
    return analytic_solution_c ? analytic_solution_c(x, y) : analytic_solution(x, y);
}

static int order_of(enum Order o) { return o == sixth ? 6 : o == fourth ? 4 : 2; }

void rhs(System sys) {
    const int Nx = sys.lat.Nx, Ny = sys.lat.Ny, X = Nx + 2, Y = Ny + 2;
    const double x0 = sys.lat.x0, y0 = sys.lat.y0, hx = sys.lat.hx, hy = sys.lat.hy;
    if (!k || !k2_y || !f || !f_y || !f_xx || !f_yy || (!analytic_solution && !analytic_solution_c)) {
        g_status = HH_ERR_ARG;
        fprintf(stderr, "ERROR: rhs: the program defines no user-input functions (k, f, ...)\n");
        return;
    }
    size_t E = (size_t)X * Y, R = 2 * (size_t)X + 2 * (size_t)Y;
    double _Complex *fe = malloc(sizeof(double _Complex) * E), *ue = malloc(sizeof(double _Complex) * E);
    double _Complex *fy = malloc(sizeof(double _Complex) * 2 * Nx);
    double _Complex *fxx = malloc(sizeof(double _Complex) * R), *fyy = malloc(sizeof(double _Complex) * R);
    double _Complex *k2e = malloc(sizeof(double _Complex) * Y), k2yb[2];
    if (!fe || !ue || !fy || !fxx || !fyy || !k2e) {
        free(fe); free(ue); free(fy); free(fxx); free(fyy); free(k2e);
        report("rhs", HH_ERR_NOMEM);
        return;
    }
    for (int j = 0; j < Y; j++) {                                      /* rhs.c:58-66 */
        double y = y0 + j * hy;
        double _Complex kk = k(y);
        k2e[j] = kk * kk;
        for (int i = 0; i < X; i++) {
            fe[i + (size_t)j * X] = f(x0 + i * hx, y);
            ue[i + (size_t)j * X] = u_at(x0 + i * hx, y);
        }
    }
    for (int i = 0; i < Nx; i++) {                                     /* 4th_order_first_derivative.c:48-50 */
        double x = x0 + (i + 1) * hx;
        fy[i] = f_y(x, y0);
        fy[Nx + i] = f_y(x, sys.lat.y1);
    }
    for (int i = 0; i < X; i++) {                                      /* ghost rings, 4th_order_second_derivative.c:41-51 */
        double x = x0 + i * hx;
        fxx[i] = f_xx(x, y0); fxx[X + i] = f_xx(x, sys.lat.y1);
        fyy[i] = f_yy(x, y0); fyy[X + i] = f_yy(x, sys.lat.y1);
    }
    for (int j = 0; j < Y; j++) {
        double y = y0 + j * hy;
        fxx[2 * X + j] = f_xx(x0, y); fxx[2 * X + Y + j] = f_xx(sys.lat.x1, y);
        fyy[2 * X + j] = f_yy(x0, y); fyy[2 * X + Y + j] = f_yy(sys.lat.x1, y);
    }
    k2yb[0] = k2_y(y0);
    k2yb[1] = k2_y(sys.lat.y1);
    for (int j = 0; j < Ny; j++) {                                     /* rhs.c:107-120 */
        double y = y0 + (j + 1) * hy;
        sys.k[j] = k(y);
        for (int i = 0; i < Nx; i++) sys.sol_analytic[i + (size_t)j * Nx] = ue[(i + 1) + (size_t)(j + 1) * X];
    }
    pthread_mutex_lock(&g_mu);
    int rc = context();
    if (!rc) rc = plan_for(&sys);
    if (!rc)
        rc = hh_rhs2d_host(g_plan, order_of(sys.order), hx, hy, (const double *)fe, (const double *)ue,
                           (const double *)fy, (const double *)fxx, (const double *)fyy, (const double *)k2e,
                           (const double *)k2yb, (double *)sys.rhs);
    pthread_mutex_unlock(&g_mu);
    free(fe); free(ue); free(fy); free(fxx); free(fyy); free(k2e);
    report("rhs", rc);
}

void residual(System sys) {
    pthread_mutex_lock(&g_mu);
    int rc = context();
    if (!rc) rc = plan_for(&sys);
    if (!rc) rc = hh_residual_host(g_plan, (const double *)sys.sol, (const double *)sys.rhs, (double *)sys.res, 1);
    pthread_mutex_unlock(&g_mu);
    if (!rc)                                                           /* residual.c:31 */
        for (size_t i = 0; i < (size_t)sys.lat.Nx * sys.lat.Ny; i++) sys.error[i] = sys.sol[i] - sys.sol_analytic[i];
    report("residual", rc);
}
