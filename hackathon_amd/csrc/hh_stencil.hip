// Stencil kernels for gfx950: compact right-hand-side assembly, the FP64
// compact-operator residual, norms and small field operations.
//
// Right-hand side (kernel set "c"):
//   2D_Program/main_functions/rhs.c:4-150 with 2D_Program/partial_derivatives
//   3D_Program/main_functions/rhs.c:13-294 with 3D_Program/partial_derivatives
// The compact fourth-order derivatives (tridiagonal [1,4,1] and [1,10,1]
// systems) run on the partitioned tridiagonal kernel (launch_compact); the
// second-order differences and the final combination are fused into one
// pointwise pass, which also subtracts the Dirichlet closure.
// Residual: 2D_Program/main_functions/residual.c:3-34 (9-point),
//           3D_Program/main_functions/residual.c:42-96 (27-point).
#include <type_traits>

#include "hh_common.h"
#include "hh_internal.h"

namespace hh {

typedef double2 z2;

HH_DEV z2 zr(double r) { return mk<double>(r, 0); }
HH_DEV z2 zs(z2 a, double s) { return cscale<double>(a, s); }
HH_DEV z2 za(z2 a, z2 b) { return cadd<double>(a, b); }
HH_DEV z2 zd(z2 a, z2 b) { return csub<double>(a, b); }
HH_DEV z2 zm(z2 a, z2 b) { return cmul<double>(a, b); }

// ---------------------------------------------------------------- 2D RHS
__global__ void k_ring2d(int Nx, int Ny, int batch, const z2 *fxx_b, const z2 *fyy_b, z2 *work, long wstride) {
    const int X = Nx + 2, Y = Ny + 2;
    const long XY = (long)X * Y;
    const int nring = 2 * X + 2 * Y;
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)nring * batch) return;
    int b = (int)(id / nring), r = (int)(id % nring);
    z2 *fxx = work + b * wstride, *fyy = fxx + XY;
    const z2 *rx = fxx_b + (long)b * nring, *ry = fyy_b + (long)b * nring;
    long o;
    if (r < X) o = r;                                   // bottom row j = 0
    else if (r < 2 * X) o = (r - X) + (long)(Y - 1) * X;     // top row j = Ny+1
    else if (r < 2 * X + Y) o = (long)(r - 2 * X) * X;       // left column i = 0
    else o = (X - 1) + (long)(r - 2 * X - Y) * X;            // right column i = Nx+1
    // rows first, then columns (2D_Program/partial_derivatives/
    // 4th_order_second_derivative.c:41-51); corners coincide analytically
    if (r < 2 * X && ((r % X) == 0 || (r % X) == X - 1)) return;
    fxx[o] = rx[r];
    fyy[o] = ry[r];
}

struct Rhs2DArgs {
    int order, Nx, Ny;
    double hx, hy;
    const z2 *f_e, *u_e, *k2c, *k2y, *coef;
    int nop;
    const z2 *work;
    long wstride;
};

__global__ void k_rhs2d_combine(Rhs2DArgs a, long total, z2 *F) {
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    const int Nx = a.Nx, Ny = a.Ny, X = Nx + 2;
    const long Nxy = (long)Nx * Ny, XY = (long)X * (Ny + 2);
    const int b = (int)(id / Nxy);
    const long p = id % Nxy;
    const int i = (int)(p % Nx), j = (int)(p / Nx);
    const z2 *f = a.f_e + b * XY;
    const long c0 = (i + 1) + (long)(j + 1) * X;
    const double hy2 = a.hy * a.hy, hy4 = hy2 * hy2, hx2 = a.hx * a.hx;
    z2 v;
    if (a.order == 6) {                                           // rhs.c:78-86
        const z2 *fxx = a.work + b * a.wstride, *fyy = fxx + XY, *fy = fyy + XY;
        z2 fxxxx = zs(za(zd(fxx[c0 - 1], zs(fxx[c0], 2.0)), fxx[c0 + 1]), 1.0 / hx2);
        z2 fxxyy = zs(za(zd(fxx[c0 - X], zs(fxx[c0], 2.0)), fxx[c0 + X]), 1.0 / hy2);
        z2 fyyyy = zs(za(zd(fyy[c0 - X], zs(fyy[c0], 2.0)), fyy[c0 + X]), 1.0 / hy2);
        const long oj = (long)(b % a.nop) * Ny + j;
        z2 t1 = zm(zd(zr(1.0), zs(a.k2c[oj], hy2 / 20.0)), f[c0]);
        z2 t2 = zs(za(fxx[c0], fyy[c0]), hy2 / 12.0);
        z2 t3 = zs(za(fxxxx, fyyyy), hy4 / 360.0);
        z2 t4 = zs(fxxyy, hy4 / 90.0);
        z2 t5 = zs(zm(a.k2y[oj], fy[p]), hy4 / 60.0);
        v = zs(za(za(za(za(t1, t2), t3), t4), t5), hy2);
    } else if (a.order == 4) {                                    // rhs.c:21-38
        z2 s4 = za(za(za(f[c0 - X], f[c0 + X]), f[c0 - 1]), f[c0 + 1]);
        v = zs(za(zs(f[c0], 2.0 / 3.0), zs(s4, 1.0 / 12.0)), hy2);
    } else {                                                      // rhs.c:4-19
        v = zs(f[c0], hy2);
    }
    // Dirichlet closure rhs.c:126-148: subtract the stencil weight times u
    // for every neighbour that lies on the boundary
    if (i == 0 || i == Nx - 1 || j == 0 || j == Ny - 1) {
        const z2 *u = a.u_e + b * XY;
        const z2 *cq = a.coef + 6 * ((long)(b % a.nop) * Ny + j);  // bm, am, b, a, bp, ap
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                int ii = i + di, jj = j + dj;
                if (ii >= 0 && ii < Nx && jj >= 0 && jj < Ny) continue;
                z2 w = dj == 1 ? (di == 0 ? cq[5] : cq[4]) : dj == 0 ? (di == 0 ? cq[3] : cq[2]) : (di == 0 ? cq[1] : cq[0]);
                v = zd(v, zm(w, u[(ii + 1) + (long)(jj + 1) * X]));
            }
    }
    F[id] = v;
}

hipError_t launch_rhs2d(const Rhs2D &r, void *F, hipStream_t st) {
    const int Nx = r.Nx, Ny = r.Ny, X = Nx + 2, Y = Ny + 2;
    const long XY = (long)X * Y, wstride = 2 * XY + (long)Nx * Ny;
    const int nring = 2 * X + 2 * Y;
    z2 *work = (z2 *)r.work;
    hipError_t e;
    if (r.order == 6) {
        // fyy: modes i along y, [1,10,1]  (4th_order_second_derivative.c:64-122)
        Compact c{};
        c.n = Ny; c.first = 0; c.diag = 10.0; c.h = r.hy;
        c.f = r.f_e; c.field_in = XY; c.mA = Nx; c.orig = 1; c.sA = 1; c.sB = 0; c.eS = X;
        c.lo = r.fyy_b; c.hi = (const z2 *)r.fyy_b + X; c.field_bnd = nring; c.bo = 1; c.bA = 1; c.bB = 0;
        c.out = work + XY; c.field_out = wstride; c.oorig = 1 + X; c.oA = 1; c.oB = 0; c.oS = X;
        c.nmodes = Nx; c.batch = r.batch;
        if ((e = launch_compact(c, st)) != hipSuccess) return e;
        // fy: modes i along y, [1,4,1]  (4th_order_first_derivative.c:36-74)
        c.first = 1; c.diag = 4.0;
        c.lo = r.fy_b; c.hi = (const z2 *)r.fy_b + Nx; c.field_bnd = 2 * Nx; c.bo = 0; c.bA = 1; c.bB = 0;
        c.out = work + 2 * XY; c.oorig = 0; c.oA = 1; c.oB = 0; c.oS = Nx;
        if ((e = launch_compact(c, st)) != hipSuccess) return e;
        // fxx: modes j along x, [1,10,1]  (4th_order_second_derivative.c:4-62)
        Compact d{};
        d.n = Nx; d.first = 0; d.diag = 10.0; d.h = r.hx;
        d.f = r.f_e; d.field_in = XY; d.mA = Ny; d.orig = X; d.sA = X; d.sB = 0; d.eS = 1;
        d.lo = (const z2 *)r.fxx_b + 2 * X; d.hi = (const z2 *)r.fxx_b + 2 * X + Y; d.field_bnd = nring;
        d.bo = 1; d.bA = 1; d.bB = 0;
        d.out = work; d.field_out = wstride; d.oorig = X + 1; d.oA = X; d.oB = 0; d.oS = 1;
        d.nmodes = Ny; d.batch = r.batch;
        if ((e = launch_compact(d, st)) != hipSuccess) return e;
        long nr = (long)nring * r.batch;
        hipLaunchKernelGGL(k_ring2d, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, st, Nx, Ny, r.batch,
                           (const z2 *)r.fxx_b, (const z2 *)r.fyy_b, work, wstride);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    Rhs2DArgs a;
    a.order = r.order; a.Nx = Nx; a.Ny = Ny; a.hx = r.hx; a.hy = r.hy;
    a.f_e = (const z2 *)r.f_e; a.u_e = (const z2 *)r.u_e; a.k2c = (const z2 *)r.k2c; a.k2y = (const z2 *)r.k2y;
    a.coef = (const z2 *)r.coef; a.nop = r.nop > 0 ? r.nop : 1; a.work = work; a.wstride = wstride;
    long total = (long)Nx * Ny * r.batch;
    hipLaunchKernelGGL(k_rhs2d_combine, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a, total, (z2 *)F);
    return hipGetLastError();
}

// ---------------------------------------------------------------- 3D RHS
__global__ void k_fill0(long n, z2 *p) {
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) p[id] = zr(0.0);
}

// ghost faces of the extended second derivatives: x faces of fxx, y faces
// of fyy, z faces of fzz (rhs.c:90-109, 4th_order_partials.c:141-146,
// 208-213, 275-280); all other ghosts stay zero
__global__ void k_faces3d(int Nx, int Ny, int Nz, int batch, const z2 *fxx_f, const z2 *fyy_f, const z2 *fzz_f,
                          z2 *work, long wstride) {
    const long X = Nx + 2, Y = Ny + 2, Z = Nz + 2, XY = X * Y, E = XY * Z;
    const long nf = 2 * Y * Z + 2 * X * Z + 2 * X * Y;
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nf * batch) return;
    int b = (int)(id / nf);
    long r = id % nf;
    z2 *w = work + b * wstride;
    if (r < 2 * Y * Z) {
        long face = r / (Y * Z), q = r % (Y * Z), j = q % Y, l = q / Y;
        w[face * (X - 1) + j * X + l * XY] = fxx_f[b * 2 * Y * Z + r];
        return;
    }
    r -= 2 * Y * Z;
    if (r < 2 * X * Z) {
        long face = r / (X * Z), q = r % (X * Z), i = q % X, l = q / X;
        w[E + i + face * (Y - 1) * X + l * XY] = fyy_f[b * 2 * X * Z + r];
        return;
    }
    r -= 2 * X * Z;
    long face = r / XY, q = r % XY;
    w[2 * E + q + face * (Z - 1) * XY] = fzz_f[b * 2 * XY + r];
}

struct Rhs3DArgs {
    int order, Nx, Ny, Nz;
    double hx, hy, hz;
    const z2 *f_e, *u_e, *k2e, *kz, *coef;
    int nop;
    const z2 *work;
    long wstride;
};

__global__ void k_rhs3d_combine(Rhs3DArgs a, long total, z2 *F) {
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    const long Nx = a.Nx, Ny = a.Ny, Nz = a.Nz, X = Nx + 2, XY = X * (Ny + 2), E = XY * (Nz + 2);
    const long N3 = Nx * Ny * Nz;
    const int b = (int)(id / N3);
    const long p = id % N3;
    const long i = p % Nx, j = (p / Nx) % Ny, l = p / (Nx * Ny);
    const z2 *f = a.f_e + b * E;
    const long c0 = (i + 1) + (j + 1) * X + (l + 1) * XY;
    const double hz = a.hz, h2 = hz * hz;
    z2 v;
    if (a.order == 6) {                                           // rhs.c:150-166
        const z2 *fxx = a.work + b * a.wstride, *fyy = fxx + E, *fzz = fyy + E;
        z2 fxxxx = zs(za(zd(fxx[c0 - 1], zs(fxx[c0], 2.0)), fxx[c0 + 1]), 1.0 / (a.hx * a.hx));
        z2 fyyyy = zs(za(zd(fyy[c0 - X], zs(fyy[c0], 2.0)), fyy[c0 + X]), 1.0 / (a.hy * a.hy));
        z2 fzzzz = zs(za(zd(fzz[c0 - XY], zs(fzz[c0], 2.0)), fzz[c0 + XY]), 1.0 / h2);
        z2 fxxyy = zs(za(zd(fxx[c0 - X], zs(fxx[c0], 2.0)), fxx[c0 + X]), 1.0 / (a.hy * a.hy));
        z2 fxxzz = zs(za(zd(fxx[c0 - XY], zs(fxx[c0], 2.0)), fxx[c0 + XY]), 1.0 / h2);
        z2 fyyzz = zs(za(zd(fyy[c0 - XY], zs(fyy[c0], 2.0)), fyy[c0 + XY]), 1.0 / h2);
        const long op = b % a.nop;
        z2 first = zm(zd(zr(1.0), zs(a.k2e[op * (Nz + 2) + l + 1], h2 / 20.0)), f[c0]);
        z2 second = zs(za(za(fxx[c0], fyy[c0]), fzz[c0]), h2 / 12.0);
        z2 third = zs(za(za(fxxxx, fyyyy), fzzzz), h2 * h2 / 360.0);
        z2 fourth = zs(za(za(fxxyy, fxxzz), fyyzz), h2 * h2 / 90.0);
        z2 fifth = zs(zm(a.kz[op * Nz + l], zd(f[c0 + XY], f[c0 - XY])), h2 * hz / 120.0);
        v = za(za(za(za(first, second), third), fourth), fifth);
    } else if (a.order == 4) {                                    // rhs.c:34-53
        z2 s6 = za(za(za(za(za(f[c0 - 1], f[c0 + 1]), f[c0 - X]), f[c0 + X]), f[c0 - XY]), f[c0 + XY]);
        v = za(f[c0], zs(zd(s6, zs(f[c0], 6.0)), 1.0 / 12.0));
    } else {                                                      // rhs.c:13-32
        v = f[c0];
    }
    // z-face closure (rhs.c:24-29, 55-62, 167-174)
    if (l == 0 || l == Nz - 1) {
        const z2 *u = a.u_e + b * E;
        for (int side = 0; side < 2; side++) {
            if ((side == 0 && l != 0) || (side == 1 && l != Nz - 1)) continue;
            long g = (i + 1) + (j + 1) * X + (side == 0 ? 0 : (Nz + 1) * XY);
            if (a.order == 2) {
                v = zd(v, u[g]);
            } else {
                const z2 *cq = a.coef + 12 * ((long)(b % a.nop) * Nz + (side == 0 ? 0 : (Nz - 1))) + (side == 0 ? 8 : 4);
                z2 corner = za(za(za(u[g - 1 - X], u[g + 1 - X]), u[g - 1 + X]), u[g + 1 + X]);
                z2 xn = za(u[g - 1], u[g + 1]);
                z2 yn = za(u[g - X], u[g + X]);
                z2 w = za(za(za(zm(cq[0], corner), zm(cq[1], xn)), zm(cq[2], yn)), zm(cq[3], u[g]));
                v = zd(v, w);
            }
        }
    }
    F[id] = v;
}

hipError_t launch_rhs3d(const Rhs3D &r, void *F, hipStream_t st) {
    const long Nx = r.Nx, Ny = r.Ny, Nz = r.Nz, X = Nx + 2, Y = Ny + 2, Z = Nz + 2, XY = X * Y, E = XY * Z;
    const long wstride = 3 * E;
    z2 *work = (z2 *)r.work;
    hipError_t e;
    if (r.order == 6) {
        long nw = wstride * r.batch;
        hipLaunchKernelGGL(k_fill0, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, st, nw, work);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        long nf = (2 * Y * Z + 2 * X * Z + 2 * X * Y) * r.batch;
        hipLaunchKernelGGL(k_faces3d, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st, r.Nx, r.Ny, r.Nz, r.batch,
                           (const z2 *)r.fxx_f, (const z2 *)r.fyy_f, (const z2 *)r.fzz_f, work, wstride);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        Compact c{};
        c.first = 0; c.diag = 10.0; c.f = r.f_e; c.field_in = E; c.field_out = wstride; c.batch = r.batch;
        // fxx: modes (j, l) along x (4th_order_partials.c:94-160)
        c.n = r.Nx; c.h = r.hx; c.mA = Ny; c.orig = X + XY; c.sA = X; c.sB = XY; c.eS = 1;
        c.lo = r.fxx_f; c.hi = (const z2 *)r.fxx_f + Y * Z; c.field_bnd = 2 * Y * Z; c.bo = 1 + Y; c.bA = 1; c.bB = Y;
        c.out = work; c.oorig = X + XY + 1; c.oA = X; c.oB = XY; c.oS = 1; c.nmodes = Ny * Nz;
        if ((e = launch_compact(c, st)) != hipSuccess) return e;
        // fyy: modes (i, l) along y (:162-227)
        c.n = r.Ny; c.h = r.hy; c.mA = Nx; c.orig = 1 + XY; c.sA = 1; c.sB = XY; c.eS = X;
        c.lo = r.fyy_f; c.hi = (const z2 *)r.fyy_f + X * Z; c.field_bnd = 2 * X * Z; c.bo = 1 + X; c.bA = 1; c.bB = X;
        c.out = work + E; c.oorig = 1 + X + XY; c.oA = 1; c.oB = XY; c.oS = X; c.nmodes = Nx * Nz;
        if ((e = launch_compact(c, st)) != hipSuccess) return e;
        // fzz: modes (i, j) along z (:229-294)
        c.n = r.Nz; c.h = r.hz; c.mA = Nx; c.orig = 1 + X; c.sA = 1; c.sB = X; c.eS = XY;
        c.lo = r.fzz_f; c.hi = (const z2 *)r.fzz_f + XY; c.field_bnd = 2 * XY; c.bo = 1 + X; c.bA = 1; c.bB = X;
        c.out = work + 2 * E; c.oorig = 1 + X + XY; c.oA = 1; c.oB = X; c.oS = XY; c.nmodes = Nx * Ny;
        if ((e = launch_compact(c, st)) != hipSuccess) return e;
    }
    Rhs3DArgs a;
    a.order = r.order; a.Nx = r.Nx; a.Ny = r.Ny; a.Nz = r.Nz; a.hx = r.hx; a.hy = r.hy; a.hz = r.hz;
    a.f_e = (const z2 *)r.f_e; a.u_e = (const z2 *)r.u_e; a.k2e = (const z2 *)r.k2e; a.kz = (const z2 *)r.kz;
    a.coef = (const z2 *)r.coef; a.nop = r.nop > 0 ? r.nop : 1; a.work = work; a.wstride = wstride;
    long total = Nx * Ny * Nz * r.batch;
    hipLaunchKernelGGL(k_rhs3d_combine, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a, total, (z2 *)F);
    return hipGetLastError();
}

// ---------------------------------------------------------------- residual
// res = A u - F with the ghosts of u zero (2D_Program/main_functions/
// residual.c:3-34, 3D_Program/main_functions/residual.c:42-96).  Each thread
// marches along the last axis over a chunk of rows (2D) or layers (3D),
// keeping the 3 x 3 (x 3) neighbourhood in registers so that every value of
// u is loaded about once per thread row; the coefficient row of a step is
// the same for the whole block.  The sums keep the order of the reference.
constexpr int RES_ROWS = 8;   // rows (2D) or layers (3D) per thread

template <typename To>
__global__ __launch_bounds__(256) void k_residual2d(int Nx, int Ny, int nchunk, const z2 *coef, int nop, const z2 *u,
                                                    const z2 *F, C2<To> *res, int b0) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j0 = blockIdx.y * RES_ROWS;
    const int b = b0 + (int)blockIdx.z;
    if (i >= Nx) return;
    (void)nchunk;
    const long Nxy = (long)Nx * Ny;
    const z2 *ub = u + b * Nxy;
    const z2 *cb = coef + 6L * (b % nop) * Ny;   // per row: bm, am, b, a, bp, ap
    auto row = [&](int jj, z2 &l, z2 &c, z2 &r) {
        if (jj < 0 || jj >= Ny) { l = c = r = zr(0.0); return; }
        const z2 *p = ub + (long)jj * Nx;
        c = p[i];
        l = i > 0 ? p[i - 1] : zr(0.0);
        r = i + 1 < Nx ? p[i + 1] : zr(0.0);
    };
    z2 lm, cm, rm, l0, c0, r0, lp, cp, rp;
    row(j0 - 1, lm, cm, rm);
    row(j0, l0, c0, r0);
    const int j1 = min(j0 + RES_ROWS, Ny);
    for (int j = j0; j < j1; j++) {
        row(j + 1, lp, cp, rp);
        const z2 *cq = cb + 6L * j;
        const z2 top = za(za(zm(cq[4], lp), zm(cq[5], cp)), zm(cq[4], rp));
        const z2 mid = za(za(zm(cq[2], l0), zm(cq[3], c0)), zm(cq[2], r0));
        const z2 bot = za(za(zm(cq[0], lm), zm(cq[1], cm)), zm(cq[0], rm));
        const long id = b * Nxy + (long)j * Nx + i;
        res[id] = cconv<To, double>(zd(za(za(top, mid), bot), F[id]));
        lm = l0; cm = c0; rm = r0;
        l0 = lp; c0 = cp; r0 = rp;
    }
}


// Tile of the 3D residual: RX x RY points of a layer per block, marching
// RZ layers along z.  Each layer's tile with its one-point halo goes
// through LDS once, loaded coalesced (three layer buffers in rotation); a
// thread reads the new layer's nine values from LDS per step and keeps, for
// each of its three layers, the four sums the stencil weights.  The loads are software-
// pipelined: a step stores the layer fetched during the previous step and
// issues the fetch of the next one (and of the next F value) before its
// barrier, so HBM latency is not paid once per layer.
constexpr int RX = 64, RY = 8, RZ = 16;

template <typename To>
__global__ __launch_bounds__(RX * RY) void k_residual3d(int Nx, int Ny, int Nz, int nchunk, const z2 *coef, int nop,
                                                       const z2 *u, const z2 *F, C2<To> *res, int b0) {
    constexpr int HX = RX + 2, HY = RY + 2, HE = HX * HY, NT = RX * RY;
    static_assert(HE <= 2 * NT, "a layer tile with halo takes at most two loads per thread");
    __shared__ z2 lay[3][HE];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * RX + tx;
    const int i0 = blockIdx.x * RX, j0 = blockIdx.y * RY;
    const int i = i0 + tx, j = j0 + ty;
    const int zc = (int)blockIdx.z % nchunk;
    const int b = b0 + (int)blockIdx.z / nchunk;
    const long Nxy = (long)Nx * Ny, N3 = Nxy * Nz;
    const z2 *ub = u + b * N3;
    // the (at most two) halo-tile elements this thread loads; points outside
    // the grid read zero (the reference's zero ghosts)
    const int e1 = tid + NT;
    const bool has1 = e1 < HE;
    auto src = [&](int e, int ll, bool &in) -> long {
        const int hy = e / HX, hx = e - hy * HX;
        const int ii = i0 + hx - 1, jj = j0 + hy - 1;
        in = ll >= 0 && ll < Nz && ii >= 0 && ii < Nx && jj >= 0 && jj < Ny;
        return in ? ii + (long)jj * Nx + (long)ll * Nxy : 0;
    };
    z2 n0, n1;   // the next layer's elements, in flight during a step
    auto fetch = [&](int ll) {
        bool in0, in1;
        const long s0 = src(tid, ll, in0), s1 = src(has1 ? e1 : tid, ll, in1);
        const z2 v0 = ub[s0], v1 = ub[s1];
        n0 = in0 ? v0 : zr(0.0);
        n1 = in1 ? v1 : zr(0.0);
    };
    auto stash = [&](int slot) {
        lay[slot][tid] = n0;
        if (has1) lay[slot][e1] = n1;
    };
    // a layer's 3 x 3 neighbourhood enters only through four coefficient-free
    // sums (corners, x and y neighbours, centre), kept instead of the nine values
    struct Part { z2 corner, xn, yn, ctr; };
    auto plane = [&](int slot, Part &p) {
        const z2 *q = &lay[slot][(ty + 1) * HX + tx + 1];
        p.corner = za(za(za(q[HX + 1], q[HX - 1]), q[-HX + 1]), q[-HX - 1]);
        p.xn = za(q[1], q[-1]);
        p.yn = za(q[HX], q[-HX]);
        p.ctr = q[0];
    };
    // one layer's contribution, coefficients c = (a, b, c, d) of that layer set
    auto layer = [&](const Part &p, const z2 *c) -> z2 {
        return za(za(za(zm(c[3], p.ctr), zm(c[0], p.corner)), zm(c[1], p.xn)), zm(c[2], p.yn));
    };
    const int l0 = zc * RZ, l1 = min(l0 + RZ, Nz);
    const bool out = i < Nx && j < Ny;
    const long id0 = b * N3 + (long)(out ? j : 0) * Nx + (out ? i : 0);
    fetch(l0 - 1);
    stash(0);
    fetch(l0);
    stash(1);
    fetch(l0 + 1);
    z2 fn = F[id0 + (long)l0 * Nxy];
    __syncthreads();
    Part pm, pc, pp;
    plane(0, pm);
    plane(1, pc);
    for (int l = l0; l < l1; l++) {
        const int sn = (l - l0 + 2) % 3;   // the slot of layer l-2, last read a barrier ago
        stash(sn);                         // layer l+1
        const z2 f = fn;
        if (l + 1 < l1) {
            fetch(l + 2);
            fn = F[id0 + (long)(l + 1) * Nxy];
        }
        __syncthreads();
        plane(sn, pp);
        if (out) {
            const z2 *cl = coef + 12 * ((long)(b % nop) * Nz + l);   // a..d, ap..dp, am..dm
            z2 acc = layer(pm, cl + 8);
            acc = za(acc, layer(pc, cl));
            acc = za(acc, layer(pp, cl + 4));
            res[id0 + (long)l * Nxy] = cconv<To, double>(zd(acc, f));
        }
        pm = pc;
        pc = pp;
        // slot sn is next rewritten two steps on; every thread has passed
        // the next step's barrier, after its read here, by then
    }
}

hipError_t launch_residual2d(int Nx, int Ny, int batch, const void *coef, int nop, const void *u, const void *F,
                             void *res, int res_f32, hipStream_t st) {
    const int nchunk = (Ny + RES_ROWS - 1) / RES_ROWS;
    for (int b0 = 0; b0 < batch; b0 += 65535) {   // grid z holds at most 65535 right-hand sides
        const int nb = batch - b0 < 65535 ? batch - b0 : 65535;
        dim3 g((unsigned)((Nx + 255) / 256), (unsigned)nchunk, (unsigned)nb);
        if (res_f32)
            hipLaunchKernelGGL(k_residual2d<float>, g, dim3(256), 0, st, Nx, Ny, nchunk, (const z2 *)coef, nop,
                               (const z2 *)u, (const z2 *)F, (float2 *)res, b0);
        else
            hipLaunchKernelGGL(k_residual2d<double>, g, dim3(256), 0, st, Nx, Ny, nchunk, (const z2 *)coef, nop,
                               (const z2 *)u, (const z2 *)F, (z2 *)res, b0);
    }
    return hipGetLastError();
}

hipError_t launch_residual3d(int Nx, int Ny, int Nz, int batch, const void *coef, int nop, const void *u,
                             const void *F, void *res, int res_f32, hipStream_t st) {
    const int nchunk = (Nz + RZ - 1) / RZ;
    const int per = 65535 / nchunk;   // right-hand sides per launch (grid z)
    if (per < 1) return hipErrorInvalidValue;
    for (int b0 = 0; b0 < batch; b0 += per) {
        const int nb = batch - b0 < per ? batch - b0 : per;
        dim3 g((unsigned)((Nx + RX - 1) / RX), (unsigned)((Ny + RY - 1) / RY), (unsigned)(nchunk * nb));
        if (res_f32)
            hipLaunchKernelGGL(k_residual3d<float>, g, dim3(RX, RY), 0, st, Nx, Ny, Nz, nchunk, (const z2 *)coef,
                               nop, (const z2 *)u, (const z2 *)F, (float2 *)res, b0);
        else
            hipLaunchKernelGGL(k_residual3d<double>, g, dim3(RX, RY), 0, st, Nx, Ny, Nz, nchunk, (const z2 *)coef,
                               nop, (const z2 *)u, (const z2 *)F, (z2 *)res, b0);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- small ops
template <typename Tc>
__global__ void k_update(long n, z2 *u, const C2<Tc> *c, double s) {
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) u[id] = za(u[id], zs(cconv<double, Tc>(c[id]), s));
}

hipError_t launch_axpy_update(long n, void *u, const void *c, int c_f32, double s, hipStream_t st) {
    dim3 g((unsigned)((n + 255) / 256));
    if (c_f32) hipLaunchKernelGGL(k_update<float>, g, dim3(256), 0, st, n, (z2 *)u, (const float2 *)c, s);
    else hipLaunchKernelGGL(k_update<double>, g, dim3(256), 0, st, n, (z2 *)u, (const z2 *)c, s);
    return hipGetLastError();
}

template <typename Ti, typename To>
__global__ void k_convert(long n, const C2<Ti> *in, C2<To> *out) {
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) out[id] = cconv<To, Ti>(in[id]);
}

hipError_t launch_convert(long n, const void *in, int in_f32, void *out, int out_f32, hipStream_t st) {
    dim3 g((unsigned)((n + 255) / 256));
    if (!in_f32 && out_f32) hipLaunchKernelGGL((k_convert<double, float>), g, dim3(256), 0, st, n, (const z2 *)in, (float2 *)out);
    else if (in_f32 && !out_f32) hipLaunchKernelGGL((k_convert<float, double>), g, dim3(256), 0, st, n, (const float2 *)in, (z2 *)out);
    else if (in_f32) hipLaunchKernelGGL((k_convert<float, float>), g, dim3(256), 0, st, n, (const float2 *)in, (float2 *)out);
    else hipLaunchKernelGGL((k_convert<double, double>), g, dim3(256), 0, st, n, (const z2 *)in, (z2 *)out);
    return hipGetLastError();
}

// max |x| and max |x - y| (L2 = false), or sqrt(sum |x|^2) and
// sqrt(sum |x - y|^2) (L2 = true), over n complex values, in two passes
// without atomics: each block writes the maxima (sums) of |x|^2 and
// |x - y|^2 over its share to part[blockIdx.x] and part[NORM_BLOCKS +
// blockIdx.x], then one block combines them and roots the result once (sqrt
// is correctly rounded and monotonic, so the maximum equals the maximum of
// the rooted values).  The grid and the reduction tree are fixed, so a sum
// is the same bits from run to run.  Four values per thread and step are
// loaded before any is used.
template <bool L2> HH_DEV double norm_join(double a, double b) { return L2 ? a + b : fmax(a, b); }

template <typename Tx, bool L2>
__global__ __launch_bounds__(256) void k_norms(long n, const C2<Tx> *x, const z2 *y, double *part) {
    __shared__ double s0[256], s1[256];
    double m0 = 0, m1 = 0;
    const long stride = (long)gridDim.x * blockDim.x;
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; id + 3 * stride < n; id += 4 * stride) {
        z2 a[4], c[4];
#pragma unroll
        for (int u = 0; u < 4; u++) a[u] = cconv<double, Tx>(x[id + u * stride]);
        if (y) {
#pragma unroll
            for (int u = 0; u < 4; u++) c[u] = y[id + u * stride];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            m0 = norm_join<L2>(m0, cabs2<double>(a[u]));
            if (y) m1 = norm_join<L2>(m1, cabs2<double>(zd(a[u], c[u])));
        }
    }
    for (; id < n; id += stride) {
        const z2 a = cconv<double, Tx>(x[id]);
        m0 = norm_join<L2>(m0, cabs2<double>(a));
        if (y) m1 = norm_join<L2>(m1, cabs2<double>(zd(a, y[id])));
    }
    s0[threadIdx.x] = m0;
    s1[threadIdx.x] = m1;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s0[threadIdx.x] = norm_join<L2>(s0[threadIdx.x], s0[threadIdx.x + s]);
            s1[threadIdx.x] = norm_join<L2>(s1[threadIdx.x], s1[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s0[0];
        part[NORM_BLOCKS + blockIdx.x] = s1[0];
    }
}

template <bool L2>
__global__ __launch_bounds__(256) void k_norms_final(int nb, const double *part, double *out) {
    __shared__ double s0[256], s1[256];
    double m0 = 0, m1 = 0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        m0 = norm_join<L2>(m0, part[b]);
        m1 = norm_join<L2>(m1, part[NORM_BLOCKS + b]);
    }
    s0[threadIdx.x] = m0;
    s1[threadIdx.x] = m1;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s0[threadIdx.x] = norm_join<L2>(s0[threadIdx.x], s0[threadIdx.x + s]);
            s1[threadIdx.x] = norm_join<L2>(s1[threadIdx.x], s1[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sqrt(s0[0]);
        out[1] = sqrt(s1[0]);
    }
}

hipError_t launch_norms(long n, const void *x, int x_f32, const void *y, double *out, hipStream_t st, bool l2) {
    long blocks = (n + 2047) / 2048;   // about eight values per thread
    if (blocks > NORM_BLOCKS) blocks = NORM_BLOCKS;
    if (blocks < 1) blocks = 1;
    double *part = out + 2;
    auto run = [&](auto L) -> hipError_t {
        constexpr bool B = decltype(L)::value;
        if (x_f32)
            hipLaunchKernelGGL((k_norms<float, B>), dim3((unsigned)blocks), dim3(256), 0, st, n, (const float2 *)x,
                               (const z2 *)y, part);
        else
            hipLaunchKernelGGL((k_norms<double, B>), dim3((unsigned)blocks), dim3(256), 0, st, n, (const z2 *)x,
                               (const z2 *)y, part);
        hipLaunchKernelGGL(k_norms_final<B>, dim3(1), dim3(256), 0, st, (int)blocks, (const double *)part, out);
        return hipGetLastError();
    };
    return l2 ? run(std::true_type()) : run(std::false_type());
}

}  // namespace hh
