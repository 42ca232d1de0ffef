/*
 * Oracle problem definitions (test infrastructure only; see oracle.h).
 *
 * The reference compiles its problem into user_input.c:
 *   2D_Program/user_input.c:4-58  k(y)=y^2+iy, u=sin x sin y, f and its
 *                                 y, xx, yy derivatives, domain [0.5,pi-0.5]^2
 *   3D_Program/user_input.c:5-44  k(z)=A-B sin(Cz), domain [0,pi]^3
 *   3D_Program/main_functions/rhs.c:229-252  u = sin(gamma x) sin(beta y)
 *                                 exp(-k(z)/C), f = -B(2A+C) sin(Cz) u
 * Here the problem is a run-time choice (kind + parameters); kind 1 adds
 * the manufactured sine solutions with constant wavenumber used by the
 * benchmark configurations.
 */
#include "problems.h"

#include <math.h>

/* ------------------------------------------------------------------ 2D */
void pr2_init(pr2 *P, int kind, const double *p) {
    P->kind = kind;
    if (kind == OR_REF) {
        P->x0 = 0.5; P->x1 = M_PI - 0.5; P->y0 = 0.5; P->y1 = M_PI - 0.5;
    } else {
        P->x0 = 0.0; P->x1 = 1.0; P->y0 = 0.0; P->y1 = 1.0;
        P->kc = p[0] + I * p[1];
        P->a = p[2]; P->b = p[3];
        P->amp = cexp(I * p[4]);
    }
}

ocplx pr2_k(const pr2 *P, double y) {
    if (P->kind == OR_REF) return y * y + I * y;           /* user_input.c:4-8 */
    return P->kc;
}
ocplx pr2_k2(const pr2 *P, double y) { ocplx k = pr2_k(P, y); return k * k; }
ocplx pr2_k2_y(const pr2 *P, double y) {
    if (P->kind == OR_REF) return 4.0 * y * y * y + 6.0 * I * y * y - 2.0 * y; /* :14-18 */
    return 0;
}
ocplx pr2_k2_yy(const pr2 *P, double y) {
    if (P->kind == OR_REF) return 12.0 * y * y + 12.0 * I * y - 2.0;          /* :20-24 */
    return 0;
}
ocplx pr2_u(const pr2 *P, double x, double y) {
    if (P->kind == OR_REF) return sin(x) * sin(y);                           /* :26-28 */
    return P->amp * sin(P->a * M_PI * x) * sin(P->b * M_PI * y);
}
/* sine problem: f = (k^2 - pi^2 (a^2+b^2)) u */
static ocplx pr2_sfac(const pr2 *P) {
    return P->kc * P->kc - M_PI * M_PI * (P->a * P->a + P->b * P->b);
}
ocplx pr2_f(const pr2 *P, double x, double y) {
    if (P->kind == OR_REF) return (pr2_k2(P, y) - 2.0) * sin(x) * sin(y);    /* :30-32 */
    return pr2_sfac(P) * pr2_u(P, x, y);
}
ocplx pr2_f_y(const pr2 *P, double x, double y) {
    if (P->kind == OR_REF)                                                   /* :34-36 */
        return (pr2_k2(P, y) - 2.0) * sin(x) * cos(y) + pr2_k2_y(P, y) * sin(x) * sin(y);
    return pr2_sfac(P) * P->amp * sin(P->a * M_PI * x) * P->b * M_PI * cos(P->b * M_PI * y);
}
ocplx pr2_f_xx(const pr2 *P, double x, double y) {
    if (P->kind == OR_REF) return (2.0 - pr2_k2(P, y)) * sin(x) * sin(y);    /* :38-40 */
    return -(P->a * M_PI) * (P->a * M_PI) * pr2_f(P, x, y);
}
ocplx pr2_f_yy(const pr2 *P, double x, double y) {
    if (P->kind == OR_REF)                                                   /* :42-45 */
        return (2.0 - pr2_k2(P, y) + pr2_k2_yy(P, y)) * sin(x) * sin(y) +
               2.0 * pr2_k2_y(P, y) * sin(x) * cos(y);
    return -(P->b * M_PI) * (P->b * M_PI) * pr2_f(P, x, y);
}

/* ------------------------------------------------------------------ 3D */
void pr3_init(pr3 *P, int kind, const double *p) {
    P->kind = kind;
    if (kind == OR_REF) {
        P->x0 = 0; P->x1 = M_PI; P->y0 = 0; P->y1 = M_PI; P->z0 = 0; P->z1 = M_PI;
        P->A = p[0]; P->B = p[1]; P->C = p[2]; P->gamma = p[3];
        P->beta = sqrt(P->A * P->A + P->B * P->B - P->gamma * P->gamma); /* user_input.c:36-40 */
    } else {
        P->x0 = 0; P->x1 = 1; P->y0 = 0; P->y1 = 1; P->z0 = 0; P->z1 = 1;
        P->kc = p[0] + I * p[1];
        P->a = p[2]; P->b = p[3]; P->c = p[4];
        P->amp = p[5] + I * p[6];
    }
}

ocplx pr3_k(const pr3 *P, double z) {
    if (P->kind == OR_REF) return P->A - P->B * sin(P->C * z);   /* user_input.c:14-16 */
    return P->kc;
}
ocplx pr3_k2(const pr3 *P, double z) { ocplx k = pr3_k(P, z); return k * k; }
/* d/dz k^2 and d2/dz2 k^2: 4th_order_partials.c:314-315, 359-360 give the
 * boundary values used by the reference for k = A - B sin(Cz). */
ocplx pr3_k2_z(const pr3 *P, double z) {
    if (P->kind == OR_REF)
        return -2.0 * P->B * P->C * cos(P->C * z) * (P->A - P->B * sin(P->C * z));
    return 0;
}
ocplx pr3_k2_zz(const pr3 *P, double z) {
    if (P->kind == OR_REF) {
        double s = sin(P->C * z), c = cos(P->C * z);
        return -2.0 * P->B * P->C * P->C * (P->B * s * s - P->A * s - P->B * c * c);
    }
    return 0;
}
ocplx pr3_u(const pr3 *P, double x, double y, double z) {
    if (P->kind == OR_REF)                                       /* rhs.c:229-248 */
        return sin(P->gamma * x) * sin(P->beta * y) * exp(-creal(pr3_k(P, z)) / P->C);
    return P->amp * sin(P->a * M_PI * x) * sin(P->b * M_PI * y) * sin(P->c * M_PI * z);
}
ocplx pr3_f(const pr3 *P, double x, double y, double z) {
    if (P->kind == OR_REF)                                       /* rhs.c:249 / hz^2 */
        return -P->B * (2.0 * P->A + P->C) * sin(P->C * z) * pr3_u(P, x, y, z);
    ocplx s = P->kc * P->kc - M_PI * M_PI * (P->a * P->a + P->b * P->b + P->c * P->c);
    return s * pr3_u(P, x, y, z);
}
ocplx pr3_f_xx(const pr3 *P, double x, double y, double z) {
    if (P->kind == OR_REF) return -P->gamma * P->gamma * pr3_f(P, x, y, z); /* rhs.c:92-93 */
    return -(P->a * M_PI) * (P->a * M_PI) * pr3_f(P, x, y, z);
}
ocplx pr3_f_yy(const pr3 *P, double x, double y, double z) {
    if (P->kind == OR_REF) return -P->beta * P->beta * pr3_f(P, x, y, z);   /* rhs.c:99-100 */
    return -(P->b * M_PI) * (P->b * M_PI) * pr3_f(P, x, y, z);
}
/* f_zz for the reference problem, written out from f = g(z) u with
 * g = -B(2A+C) sin(Cz), u_z = u B cos(Cz), u_zz = u (B^2 cos^2 - BC sin).
 * (rhs.c:106-107 states a value only at the z faces.) */
ocplx pr3_f_zz(const pr3 *P, double x, double y, double z) {
    if (P->kind == OR_REF) {
        double A = P->A, B = P->B, C = P->C;
        double s = sin(C * z), c = cos(C * z);
        double g = -B * (2 * A + C) * s;
        double gz = -B * (2 * A + C) * C * c;
        double gzz = B * (2 * A + C) * C * C * s;
        double uz_u = B * c;
        double uzz_u = B * B * c * c - B * C * s;
        return (gzz + 2.0 * gz * uz_u + g * uzz_u) * pr3_u(P, x, y, z);
    }
    return -(P->c * M_PI) * (P->c * M_PI) * pr3_f(P, x, y, z);
}
