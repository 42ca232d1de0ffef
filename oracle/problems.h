/* Oracle problem definitions (test infrastructure only; see oracle.h). */
#ifndef HH_ORACLE_PROBLEMS_H
#define HH_ORACLE_PROBLEMS_H
#include "oracle.h"

typedef struct {
    int kind;
    double x0, x1, y0, y1;
    ocplx kc, amp;
    double a, b;
} pr2;

void pr2_init(pr2 *P, int kind, const double *p);
ocplx pr2_k(const pr2 *P, double y);
ocplx pr2_k2(const pr2 *P, double y);
ocplx pr2_k2_y(const pr2 *P, double y);
ocplx pr2_k2_yy(const pr2 *P, double y);
ocplx pr2_u(const pr2 *P, double x, double y);
ocplx pr2_f(const pr2 *P, double x, double y);
ocplx pr2_f_y(const pr2 *P, double x, double y);
ocplx pr2_f_xx(const pr2 *P, double x, double y);
ocplx pr2_f_yy(const pr2 *P, double x, double y);

typedef struct {
    int kind;
    double x0, x1, y0, y1, z0, z1;
    double A, B, C, gamma, beta;
    ocplx kc, amp;
    double a, b, c;
} pr3;

void pr3_init(pr3 *P, int kind, const double *p);
ocplx pr3_k(const pr3 *P, double z);
ocplx pr3_k2(const pr3 *P, double z);
ocplx pr3_k2_z(const pr3 *P, double z);
ocplx pr3_k2_zz(const pr3 *P, double z);
ocplx pr3_u(const pr3 *P, double x, double y, double z);
ocplx pr3_f(const pr3 *P, double x, double y, double z);
ocplx pr3_f_xx(const pr3 *P, double x, double y, double z);
ocplx pr3_f_yy(const pr3 *P, double x, double y, double z);
ocplx pr3_f_zz(const pr3 *P, double x, double y, double z);

typedef struct or_dstplan or_dstplan;
or_dstplan *or_dst_plan(int N);
void or_dst_plan_free(or_dstplan *d);
void or_dst_lines_plan(const or_dstplan *d, long nlines, const ocplx *in, long iline,
                       long istride, ocplx *out, long oline, long ostride, int threads);
void or_dst_lines(int N, long nlines, const ocplx *in, long iline, long istride,
                  ocplx *out, long oline, long ostride, int threads);
#endif
