/*
 * Direct_Solver3d: the 3D program's driver over the drop-in GPU library.
 *
 * Mirrors 3D_Program/main_functions/main.c:4-39 and setup.c:6-108:
 *   ./Direct_Solver3d order Nx [Ny [Nz]]      (defaults: order 2, N = 100)
 * Problem (3D_Program/user_input.c:5-44): k(z) = A - B sin(Cz),
 * u = sin(gamma x) sin(beta y) exp(-k(z)/C) on [0, pi]^3 with A = 10,
 * B = 0, C = 10, gamma = 6 as shipped; HH_PROBLEM="ref:A,B,C,gamma"
 * selects other values (e.g. 10,9,10,9 for the paper's Tables 4-6).
 * background3D and coefficients run on the host; rhs, solver and residual
 * are the GPU drop-ins of libhh_ref3d.so.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hh_ref3d.h"
#include "hh_solver.h"

static double now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static double normInf(long n, const double _Complex *x) {
    double m = 0;
    for (long i = 0; i < n; i++)
        if (cabs(x[i]) > m) m = cabs(x[i]);
    return m;
}

static double normL2(long n, const double _Complex *x) {   /* helpers_functions/norms.c:15-24 */
    double s = 0;
    for (long i = 0; i < n; i++) s += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
    return sqrt(s);
}

int main(int argc, char **argv) {
    System sys;
    memset(&sys, 0, sizeof sys);
    sys.lat.x0 = 0; sys.lat.x1 = M_PI; sys.lat.y0 = 0; sys.lat.y1 = M_PI; sys.lat.z0 = 0; sys.lat.z1 = M_PI;
    sys.A = 10.0; sys.B = 0.0; sys.C = 10.0; sys.gamma = 6.0;
    const char *pr = getenv("HH_PROBLEM");
    if (pr && !strncmp(pr, "ref:", 4) && sscanf(pr + 4, "%lf,%lf,%lf,%lf", &sys.A, &sys.B, &sys.C, &sys.gamma) != 4) {
        fprintf(stderr, "HH_PROBLEM=ref:A,B,C,gamma\n");
        return 2;
    }
    sys.beta = sqrt(sys.A * sys.A + sys.B * sys.B - sys.gamma * sys.gamma);
    int order = argc > 1 ? atoi(argv[1]) : 2;
    sys.order = order == 6 ? sixth : order == 4 ? fourth : second;
    sys.lat.Nx = argc > 2 ? atoi(argv[2]) : 100;
    sys.lat.Ny = argc > 3 ? atoi(argv[3]) : sys.lat.Nx;
    sys.lat.Nz = argc > 4 ? atoi(argv[4]) : sys.lat.Nx;
    const int Nx = sys.lat.Nx, Ny = sys.lat.Ny, Nz = sys.lat.Nz;
    if (Nx < 2 || Ny < 2 || Nz < 2) { fprintf(stderr, "grid sizes must be at least 2\n"); return 2; }
    sys.lat.hx = (sys.lat.x1 - sys.lat.x0) / (Nx + 1);
    sys.lat.hy = (sys.lat.y1 - sys.lat.y0) / (Ny + 1);
    sys.lat.hz = (sys.lat.z1 - sys.lat.z0) / (Nz + 1);
    sys.lat.Nxyz = Nx * Ny * Nz;
    size_t n = (size_t)Nx * Ny * Nz, cs = sizeof(double _Complex);
    double _Complex **co[12] = {&sys.a, &sys.b, &sys.c, &sys.d, &sys.ap, &sys.bp,
                                &sys.cp, &sys.dp, &sys.am, &sys.bm, &sys.cm, &sys.dm};
    for (int q = 0; q < 12; q++) *co[q] = malloc(Nz * cs);
    sys.k_bg_ext = malloc((Nz + 2) * cs); sys.k2_bg_ext = malloc((Nz + 2) * cs);
    sys.sol_analytic = malloc(n * cs); sys.rhs = malloc(n * cs); sys.sol = malloc(n * cs);
    sys.res = malloc(n * cs); sys.error = malloc(n * cs);
    printf("\n\t%s Order\n\tGrid %d x %d x %d\n", order == 6 ? "6th" : order == 4 ? "4th" : "2nd", Nx, Ny, Nz);

    for (int l = 0; l < Nz + 2; l++) {                                    /* background3D */
        double z = sys.lat.z0 + sys.lat.hz * l;
        sys.k_bg_ext[l] = sys.A - sys.B * sin(sys.C * z);
        sys.k2_bg_ext[l] = sys.k_bg_ext[l] * sys.k_bg_ext[l];
    }
    double A = sys.A, B = sys.B, C = sys.C, z0 = sys.lat.z0, z1 = sys.lat.z1;
    double _Complex k2d[4] = {-2.0 * B * C * cos(C * z0) * (A - B * sin(C * z0)),
                              -2.0 * B * C * cos(C * z1) * (A - B * sin(C * z1)),
                              -2.0 * B * C * C * (B * sin(C * z0) * sin(C * z0) - A * sin(C * z0) - B * cos(C * z0) * cos(C * z0)),
                              -2.0 * B * C * C * (B * sin(C * z1) * sin(C * z1) - A * sin(C * z1) - B * cos(C * z1) * cos(C * z1))};
    double *cp[12];
    for (int q = 0; q < 12; q++) cp[q] = (double *)*co[q];
    int rc = hh_coefficients3d(order, Nx, Ny, Nz, sys.lat.hx, sys.lat.hy, sys.lat.hz, (double *)sys.k2_bg_ext,
                               (double *)k2d, cp);
    if (rc) { fprintf(stderr, "coefficients: %s\n", hh_last_error()); return 1; }
    rhs(sys);
    if (hh_ref3d_last_status()) return 1;

    double t0 = now();
    solver(sys);
    double t1 = now();
    if (hh_ref3d_last_status()) return 1;
    printf("\n \tSolver wall-time = %f sec (first call: device and plan set-up included)\n", t1 - t0);
    t0 = now();
    solver(sys);
    printf("   \tSolver wall-time = %f sec (plan reused)\n\n", now() - t0);

    residual(sys);
    if (hh_ref3d_last_status()) return 1;
    printf("\t||error||_inf =  %10.7e \n\n", normInf(n, sys.error));
    printf("\t||res||_inf =  %10.7e \n\n", normInf(n, sys.res));
    printf("\t||res||_2 =  %10.7e \n\n", normL2(n, sys.res));
    return 0;
}
