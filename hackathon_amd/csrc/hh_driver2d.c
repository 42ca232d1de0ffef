/*
 * Direct_Solver2d: the 2D program's driver over the drop-in GPU library.
 *
 * Mirrors 2D_Program/main_functions/main.c:4-46 and setup.c:4-71:
 *   ./Direct_Solver2d order N [threads]      e.g. ./Direct_Solver2d 6 100 2
 * (no interactive prompts: missing arguments take order 2, N = 100).
 * Problem (user input, 2D_Program/user_input.c:4-58): by default the
 * reference's k(y) = y^2 + iy, u = sin x sin y on [0.5, pi-0.5]^2.  With
 * HH_PROBLEM="sine:k,a,b,re,im" the manufactured solution
 * u = (re + i im) sin(a pi x) sin(b pi y) with constant k on the unit square.
 * coefficients run on the host (hh_coefficients2d); rhs, solver and residual
 * are the GPU drop-ins of libhh_ref2d.so.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hh_ref2d.h"
#include "hh_solver.h"

static int g_sine;
static double g_k, g_a, g_b;
static double _Complex g_amp = 1.0;

/* ---- user input (2D_Program/user_input.c) ---- */
double _Complex k(double y) { return g_sine ? g_k : y * y + I * y; }
double _Complex k2_y(double y) { return g_sine ? 0 : 4.0 * y * y * y + 6.0 * I * y * y - 2.0 * y; }
double _Complex k2_yy(double y) { return g_sine ? 0 : 12.0 * y * y + 12.0 * I * y - 2.0; }
static double _Complex sfac(void) { return g_k * g_k - M_PI * M_PI * (g_a * g_a + g_b * g_b); }
double analytic_solution(double x, double y) {
    return g_sine ? creal(g_amp) * sin(g_a * M_PI * x) * sin(g_b * M_PI * y) : sin(x) * sin(y);
}
/* complex Dirichlet data (extension: the reference's analytic_solution is real) */
double _Complex analytic_solution_c(double x, double y) {
    return g_sine ? g_amp * sin(g_a * M_PI * x) * sin(g_b * M_PI * y) : sin(x) * sin(y);
}
double _Complex f(double x, double y) {
    if (!g_sine) return (k(y) * k(y) - 2.0) * sin(x) * sin(y);
    return sfac() * analytic_solution_c(x, y);
}
double _Complex f_y(double x, double y) {
    if (!g_sine) return (k(y) * k(y) - 2.0) * sin(x) * cos(y) + k2_y(y) * sin(x) * sin(y);
    return sfac() * g_amp * sin(g_a * M_PI * x) * g_b * M_PI * cos(g_b * M_PI * y);
}
double _Complex f_xx(double x, double y) {
    if (!g_sine) return (2.0 - k(y) * k(y)) * sin(x) * sin(y);
    return -(g_a * M_PI) * (g_a * M_PI) * f(x, y);
}
double _Complex f_yy(double x, double y) {
    if (!g_sine) return (2.0 - k(y) * k(y) + k2_yy(y)) * sin(x) * sin(y) + 2.0 * k2_y(y) * sin(x) * cos(y);
    return -(g_b * M_PI) * (g_b * M_PI) * f(x, y);
}

static double now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static double normInf(long n, const double _Complex *x) {   /* helpers_functions/norms.c:3-13 */
    double m = 0;
    for (long i = 0; i < n; i++)
        if (cabs(x[i]) > m) m = cabs(x[i]);
    return m;
}

int main(int argc, char **argv) {
    System sys;
    memset(&sys, 0, sizeof sys);
    const char *pr = getenv("HH_PROBLEM");
    double re = 1, im = 0;
    if (pr && !strncmp(pr, "sine:", 5)) {
        g_sine = 1;
        if (sscanf(pr + 5, "%lf,%lf,%lf,%lf,%lf", &g_k, &g_a, &g_b, &re, &im) < 3) {
            fprintf(stderr, "HH_PROBLEM=sine:k,a,b[,re,im]\n");
            return 2;
        }
        g_amp = re + I * im;
    }
    if (g_sine) { sys.lat.x0 = 0; sys.lat.x1 = 1; sys.lat.y0 = 0; sys.lat.y1 = 1; }
    else { sys.lat.x0 = 0.5; sys.lat.x1 = M_PI - 0.5; sys.lat.y0 = 0.5; sys.lat.y1 = M_PI - 0.5; }
    int order = argc > 1 ? atoi(argv[1]) : 2;
    sys.order = order == 6 ? sixth : order == 4 ? fourth : second;
    sys.lat.Nx = argc > 2 ? atoi(argv[2]) : 100;
    sys.lat.Ny = sys.lat.Nx;
    sys.threadCount = argc > 3 ? atoi(argv[3]) : 1;
    if (sys.lat.Nx < 2) { fprintf(stderr, "N must be at least 2\n"); return 2; }
    const int Nx = sys.lat.Nx, Ny = sys.lat.Ny;
    sys.lat.hx = (sys.lat.x1 - sys.lat.x0) / (Nx + 1);
    sys.lat.hy = (sys.lat.y1 - sys.lat.y0) / (Ny + 1);
    sys.lat.Nxy = Nx * Ny;
    size_t n = (size_t)Nx * Ny, cs = sizeof(double _Complex);
    sys.sol = malloc(n * cs); sys.sol_analytic = malloc(n * cs); sys.rhs = malloc(n * cs);
    sys.res = malloc(n * cs); sys.error = malloc(n * cs); sys.k = malloc(Ny * cs);
    sys.bp = malloc(Ny * cs); sys.ap = malloc(Ny * cs); sys.b = malloc(Ny * cs);
    sys.a = malloc(Ny * cs); sys.bm = malloc(Ny * cs); sys.am = malloc(Ny * cs);

    printf("\n\t%s Order\n\tGrid %d x %d\n\tOpenMP Threads: %d\n", order == 6 ? "6th" : order == 4 ? "4th" : "2nd",
           Nx, Ny, sys.threadCount);
    double t0 = now();
    /* coefficients (coefficients.c:63-67) on the host */
    double _Complex *k2e = malloc((Ny + 2) * cs), *k2y = malloc(Ny * cs), *k2yy = malloc(Ny * cs);
    for (int j = 0; j < Ny + 2; j++) { double _Complex kk = k(sys.lat.y0 + j * sys.lat.hy); k2e[j] = kk * kk; }
    for (int j = 0; j < Ny; j++) {
        double y = sys.lat.y0 + (j + 1) * sys.lat.hy;
        k2y[j] = k2_y(y);
        k2yy[j] = k2_yy(y);
    }
    int rc = hh_coefficients2d(order, Nx, Ny, sys.lat.hx, sys.lat.hy, (double *)k2e, (double *)k2y, (double *)k2yy,
                               (double *)sys.bp, (double *)sys.ap, (double *)sys.b, (double *)sys.a,
                               (double *)sys.bm, (double *)sys.am);
    if (rc) { fprintf(stderr, "coefficients: %s\n", hh_last_error()); return 1; }
    rhs(sys);
    if (hh_ref2d_last_status()) return 1;
    printf("\n \tSet up wall-time = %f sec \n", now() - t0);

    t0 = now();
    solver(sys);
    double t1 = now();
    if (hh_ref2d_last_status()) return 1;
    printf("\n \tSolver wall-time = %f sec (first call: device and plan set-up included)\n", t1 - t0);
    t0 = now();
    solver(sys);
    printf("   \tSolver wall-time = %f sec (plan reused)\n\n", now() - t0);

    residual(sys);
    if (hh_ref2d_last_status()) return 1;
    printf("\t||res||_inf =  %10.7e \n\n", normInf(n, sys.res));
    printf("\t||error||_inf =  %10.7e \n\n", normInf(n, sys.error));
    return 0;
}
