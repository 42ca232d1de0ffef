/*
 * Oracle FFT and DST-I (test infrastructure only; see oracle.h).
 *
 * The reference computes its orthonormal DST-I with FFTW r2c transforms of
 * the zero-padded rows (2D_Program/solver_functions/DST.c:6-44,
 * 2D_Program/solver_functions/solver.c:24-25,61-62;
 * 3D_Program/solver_functions/DST.c:148-204).  FFTW is absent here, so this
 * file supplies a plain complex FFT:
 *   - iterative radix-2 decimation-in-time for power-of-two lengths;
 *   - Bluestein's chirp-z algorithm (L. Bluestein, 1970) on top of it for
 *     every other length (e.g. 2(N+1) = 202 for N = 100).
 * The DST-I of a complex row x (length N) is taken from the FFT of its odd
 * extension z = [0, x, 0, -reverse(x)] of length L = 2(N+1):
 *   Z_k = -2i sum_n x_n sin(pi n k/(N+1))   =>   X_k = sqrt(2/(N+1)) (i/2) Z_k.
 * One complex FFT of the odd extension does the work of the reference's two
 * real transforms (real and imaginary parts) of length 2N+2.
 */
#include "problems.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct fftp {
    int L;
    int pow2;        /* 1: radix-2 path, 0: Bluestein */
    ocplx *tw;       /* radix-2: exp(-2 pi i k / L), k < L/2 */
    int *rev;        /* radix-2: bit reversal permutation */
    int M;           /* Bluestein: power-of-two convolution length */
    ocplx *w;        /* Bluestein: exp(-pi i n^2 / L), n < L */
    ocplx *bhat;     /* Bluestein: FFT_M of the conjugate chirp */
    struct fftp *sub;
} fftp;

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void fft_pow2(const fftp *p, ocplx *x) {
    int L = p->L;
    for (int i = 0; i < L; i++) {
        int r = p->rev[i];
        if (r > i) { ocplx t = x[i]; x[i] = x[r]; x[r] = t; }
    }
    for (int len = 2; len <= L; len <<= 1) {
        int half = len >> 1, step = L / len;
        for (int s = 0; s < L; s += len) {
            for (int k = 0; k < half; k++) {
                ocplx w = p->tw[k * step];
                ocplx u = x[s + k], v = x[s + k + half] * w;
                x[s + k] = u + v;
                x[s + k + half] = u - v;
            }
        }
    }
}

static fftp *fft_create(int L);

static void fft_free(fftp *p) {
    if (!p) return;
    free(p->tw); free(p->rev); free(p->w); free(p->bhat);
    fft_free(p->sub);
    free(p);
}

static fftp *fft_create(int L) {
    fftp *p = (fftp *)calloc(1, sizeof(fftp));
    p->L = L;
    if (is_pow2(L)) {
        p->pow2 = 1;
        p->tw = (ocplx *)malloc(sizeof(ocplx) * (L / 2 + 1));
        for (int k = 0; k <= L / 2; k++) {
            double ang = -2.0 * M_PI * (double)k / (double)L;
            p->tw[k] = cos(ang) + I * sin(ang);
        }
        p->rev = (int *)malloc(sizeof(int) * L);
        int lg = 0;
        while ((1 << lg) < L) lg++;
        for (int i = 0; i < L; i++) {
            int r = 0;
            for (int b = 0; b < lg; b++) if (i & (1 << b)) r |= 1 << (lg - 1 - b);
            p->rev[i] = r;
        }
        return p;
    }
    p->pow2 = 0;
    int M = 1;
    while (M < 2 * L - 1) M <<= 1;
    p->M = M;
    p->sub = fft_create(M);
    p->w = (ocplx *)malloc(sizeof(ocplx) * L);
    for (long n = 0; n < L; n++) {
        long q = (n * n) % (2L * L);           /* keep the angle small */
        double ang = -M_PI * (double)q / (double)L;
        p->w[n] = cos(ang) + I * sin(ang);
    }
    p->bhat = (ocplx *)calloc(M, sizeof(ocplx));
    p->bhat[0] = conj(p->w[0]);
    for (int n = 1; n < L; n++) {
        p->bhat[n] = conj(p->w[n]);
        p->bhat[M - n] = conj(p->w[n]);
    }
    fft_pow2(p->sub, p->bhat);
    return p;
}

/* forward DFT, X_k = sum_n x_n exp(-2 pi i nk/L); work has >= M entries */
static void fft_exec(const fftp *p, ocplx *x, ocplx *work) {
    if (p->pow2) { fft_pow2(p, x); return; }
    int L = p->L, M = p->M;
    for (int n = 0; n < L; n++) work[n] = x[n] * p->w[n];
    for (int n = L; n < M; n++) work[n] = 0;
    fft_pow2(p->sub, work);
    for (int n = 0; n < M; n++) work[n] *= p->bhat[n];
    /* inverse via conjugation */
    for (int n = 0; n < M; n++) work[n] = conj(work[n]);
    fft_pow2(p->sub, work);
    double inv = 1.0 / (double)M;
    for (int k = 0; k < L; k++) x[k] = p->w[k] * conj(work[k]) * inv;
}

/* One DST-I row: odd extension, FFT, X_k = c (i/2) Z_{k+1}. */
static void dst1_row(const fftp *p, int N, const ocplx *in, long istride,
                     ocplx *out, long ostride, ocplx *z, ocplx *work) {
    int L = p->L;
    z[0] = 0;
    z[N + 1] = 0;
    for (int n = 1; n <= N; n++) {
        ocplx v = in[(long)(n - 1) * istride];
        z[n] = v;
        z[L - n] = -v;
    }
    fft_exec(p, z, work);
    double c = 0.5 * sqrt(2.0 / (double)(N + 1));
    for (int k = 1; k <= N; k++) out[(long)(k - 1) * ostride] = c * I * z[k];
}

struct or_dstplan {
    fftp *p;
    int N;
};

or_dstplan *or_dst_plan(int N) {
    or_dstplan *d = (or_dstplan *)malloc(sizeof(or_dstplan));
    d->N = N;
    d->p = fft_create(2 * (N + 1));
    return d;
}

void or_dst_plan_free(or_dstplan *d) {
    if (!d) return;
    fft_free(d->p);
    free(d);
}

/* nlines independent DSTs: line r starts at in + r*iline, elements istride apart */
void or_dst_lines_plan(const or_dstplan *d, long nlines, const ocplx *in, long iline,
                       long istride, ocplx *out, long oline, long ostride, int threads) {
    int N = d->N, L = 2 * (N + 1);
    size_t wl = (size_t)(d->p->pow2 ? d->p->L : d->p->M);
    if (threads < 1) threads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads) if (threads > 1)
#endif
    {
        ocplx *z = (ocplx *)malloc(sizeof(ocplx) * L);
        ocplx *work = (ocplx *)malloc(sizeof(ocplx) * wl);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (long r = 0; r < nlines; r++)
            dst1_row(d->p, N, in + r * iline, istride, out + r * oline, ostride, z, work);
        free(z);
        free(work);
    }
}

void or_dst_lines(int N, long nlines, const ocplx *in, long iline, long istride,
                  ocplx *out, long oline, long ostride, int threads) {
    or_dstplan *d = or_dst_plan(N);
    or_dst_lines_plan(d, nlines, in, iline, istride, out, oline, ostride, threads);
    or_dst_plan_free(d);
}

void or_dst1_rows(int N, long nrows, const ocplx *in, ocplx *out, int threads) {
    or_dst_lines(N, nrows, in, N, 1, out, N, 1, threads);
}

double or_normInf(long n, const ocplx *x) {
    double m = 0;
    for (long i = 0; i < n; i++) {
        double v = cabs(x[i]);
        if (v > m) m = v;
    }
    return m;
}

double or_normL2(long n, const ocplx *x) {
    double s = 0;
    for (long i = 0; i < n; i++) s += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
    return sqrt(s);
}

double or_wtime(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
