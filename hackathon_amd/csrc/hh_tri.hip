// Per-mode tridiagonal solves for gfx950 (the reference's "Middle stuff":
// 2D_Program/solver_functions/solver.c:71-86 with the factors of
// 2D_Program/main_functions/LU.c:3-46; 3D_Program/solver_functions/
// solver.c:356-369 / cuda_kernels.cu:208-249 with 3D_Program/main_functions/
// LU.c:4-81).
//
// Mode m owns the system  sub_q x_{q-1} + dia_q x_q + sup_q x_{q+1} = d_q,
// q = 0..n-1, along the last axis (y in 2D, z in 3D).  The bands are
// rebuilt on the fly from the O(n) stencil coefficient vectors and the
// closed-form eigenvalues 2cos(k pi/(N+1)); nothing of size N^2 or N^3 is
// stored (the reference streams 48 B/DOF of L, U, Up).
//
// Block = W neighbouring modes x P chunks.  Thread (w, c) owns rows
// [s, e] of mode w in registers (row-major layout is kept: lanes w are
// adjacent in memory, so each load instruction moves W*16-byte segments).
//   1. downward sweep: row q in the form  a~_q x_s + x_q + c~_q x_{q+1} = d~_q
//   2. upward sweep:   row q in the form  a^_q x_s + x_q + c^_q x_e = d^_q
//   3. the chunk's first and last rows give two rows of a reduced
//      tridiagonal system of size 2P in (x_s^0, x_e^0, x_s^1, ...), solved by
//      parallel cyclic reduction in LDS with normalised rows;
//   4. x_q = d^_q - a^_q x_s - c^_q x_e, store.
// HBM is read once and written once; the solve may run in place.
#include "hh_common.h"
#include "hh_internal.h"

namespace hh {

// Helmholtz mode bands.  coef is an array of complex rows: 2D [q][6] =
// (bm, am, b, a, bp, ap) with sub = bm lam + am etc.; 3D [q][12] =
// (a, b, c, d, ap, bp, cp, dp, am, bm, cm, dm) with
// dia = a lx ly + b lx + c ly + d.
// With nop > 1 operators, right-hand side b uses the table at
// coef + (b % nop) * cstride (one operator per problem of a batch).
// One row of a tridiagonal system, returned by value (reference outputs
// written under branches defeat scalar replacement and land in scratch).
template <typename T> struct Band {
    C2<T> sub, dia, sup;
};

template <typename T, int DIM> struct HelmBands {
    int n, Nx, nop;
    const C2<T> *coef;
    long cstride;
    const double *lamx, *lamy;
    struct Mode { T l0, l1, l2; const C2<T> *cf; };
    HH_DEV Mode mode(long m, int bt) const {
        Mode md;
        md.cf = coef + (long)(bt % nop) * cstride;
        if constexpr (DIM == 2) {
            md.l0 = (T)lamx[m]; md.l1 = 0; md.l2 = 0;
        } else {
            long i = m % Nx, j = m / Nx;
            T lx = (T)lamx[i], ly = (T)lamy[j];
            md.l0 = lx * ly; md.l1 = lx; md.l2 = ly;
        }
        return md;
    }
    HH_DEV Band<T> band(const Mode &md, int q) const {
        Band<T> r;
        const T ms = q > 0 ? (T)1 : (T)0, ps = q < n - 1 ? (T)1 : (T)0;
        if constexpr (DIM == 2) {
            const C2<T> *c = md.cf + 6 * (long)q;
            const T l = md.l0;
            r.sub = mk<T>(ms * (c[0].x * l + c[1].x), ms * (c[0].y * l + c[1].y));
            r.dia = mk<T>(c[2].x * l + c[3].x, c[2].y * l + c[3].y);
            r.sup = mk<T>(ps * (c[4].x * l + c[5].x), ps * (c[4].y * l + c[5].y));
        } else {
            const C2<T> *c = md.cf + 12 * (long)q;
            const T a = md.l0, b = md.l1, g = md.l2;
            r.dia = mk<T>(c[0].x * a + c[1].x * b + c[2].x * g + c[3].x, c[0].y * a + c[1].y * b + c[2].y * g + c[3].y);
            r.sup = mk<T>(ps * (c[4].x * a + c[5].x * b + c[6].x * g + c[7].x),
                          ps * (c[4].y * a + c[5].y * b + c[6].y * g + c[7].y));
            r.sub = mk<T>(ms * (c[8].x * a + c[9].x * b + c[10].x * g + c[11].x),
                          ms * (c[8].y * a + c[9].y * b + c[10].y * g + c[11].y));
        }
        return r;
    }
};

// Helmholtz problem: bands from HelmBands, right-hand side and solution in
// the row-major field (element (b, m, q) at b*field + m + q*S).
// the row-major field (element (b, m, q) at b*field + m + q*S), or, with
// tw > 0, the mode-tiled one the row DSTs write for it (tiles of tw
// neighbouring modes, tile t0 = m & ~(tw-1) holding its modes for all n rows
// contiguously: b*field + t0*n + q*w + (m - t0), w = min(tw, nmodes - t0)),
// where each thread's rows are whole w*16-byte segments one after another.
template <typename T, typename Ti, typename To, int DIM> struct HelmProb {
    HelmBands<T, DIM> bands;
    const C2<Ti> *in;
    C2<To> *out;
    long field, S, nmodes;
    int tw;
    struct Mode { typename HelmBands<T, DIM>::Mode bm; long off, step; };
    HH_DEV int n() const { return bands.n; }
    HH_DEV Mode mode(long m, int bt) const {
        Mode md;
        md.bm = bands.mode(m, bt);
        if (tw == 0) {
            md.off = m;
            md.step = S;
        } else {
            const long t0 = m & ~(long)(tw - 1);
            md.step = min((long)tw, nmodes - t0);
            md.off = t0 * bands.n + (m - t0);
        }
        return md;
    }
    HH_DEV Band<T> band(const Mode &md, int q) const { return bands.band(md.bm, q); }
    HH_DEV C2<T> rhs(int b, const Mode &md, int q) const {
        return cconv<T, Ti>(in[b * field + md.off + (long)q * md.step]);
    }
    HH_DEV void store(int b, const Mode &md, int q, C2<T> x) const {
        out[b * field + md.off + (long)q * md.step] = cconv<To, T>(x);
    }
};

// Compact fourth-order derivative of a field sampled on the extended grid
// (2D_Program/partial_derivatives/4th_order_*.c, 3D_Program/
// partial_derivatives/4th_order_partials.c:94-294): band (1, diag, 1),
//   second derivative (diag 10): rhs = 12 (f_{q+1} - 2 f_q + f_{q-1}) / h^2
//   first derivative  (diag 4):  rhs = 3 (f_{q+1} - f_{q-1}) / h
// minus the analytic boundary value at q = 0 and q = n-1.  Mode m maps to
// the extended-grid line origin + (m % mA) sA + (m / mA) sB; element q sits
// at (q+1) eS along it.  The result goes to oorig + (m%mA) oA + (m/mA) oB + q oS.
template <typename T> struct CompactProb {
    int nn, first;
    T diag, h;
    const C2<T> *f;
    long field_in, field_out, field_bnd;
    long mA, orig, sA, sB, eS;
    const C2<T> *lo, *hi;   // boundary values, indexed bo + (m%mA) bA + (m/mA) bB
    long bo, bA, bB;
    C2<T> *out;
    long oorig, oA, oB, oS;
    struct Mode { long in, bnd, out; };
    HH_DEV int n() const { return nn; }
    HH_DEV Mode mode(long m, int) const {
        long a = m % mA, b = m / mA;
        Mode md;
        md.in = orig + a * sA + b * sB;
        md.bnd = bo + a * bA + b * bB;
        md.out = oorig + a * oA + b * oB;
        return md;
    }
    HH_DEV Band<T> band(const Mode &, int q) const {
        Band<T> r;
        r.sub = mk<T>(q > 0 ? (T)1 : (T)0, 0);
        r.dia = mk<T>(diag, 0);
        r.sup = mk<T>(q < nn - 1 ? (T)1 : (T)0, 0);
        return r;
    }
    HH_DEV C2<T> rhs(int bt, const Mode &md, int q) const {
        const C2<T> *g = f + bt * field_in + md.in + (long)(q + 1) * eS;
        C2<T> v;
        if (first) v = cscale<T>(csub<T>(g[eS], g[-eS]), (T)3 / h);
        else v = cscale<T>(cadd<T>(csub<T>(g[eS], cscale<T>(g[0], (T)2)), g[-eS]), (T)12 / (h * h));
        if (q == 0) v = csub<T>(v, lo[bt * field_bnd + md.bnd]);
        if (q == nn - 1) v = csub<T>(v, hi[bt * field_bnd + md.bnd]);
        return v;
    }
    HH_DEV void store(int bt, const Mode &md, int q, C2<T> x) const {
        out[bt * field_out + md.out + (long)q * oS] = x;
    }
};

// Blocks are dealt round-robin to the 8 XCDs (block b runs on XCD b % 8).
// Neighbouring mode tiles share cache lines when W*sizeof(complex) < 128,
// so tiles are renumbered to run consecutively on one XCD and its L2.
HH_DEV long xcd_tile(long b, long G) {
    const long G8 = G & ~7L;
    if (b >= G8) return b;
    return (b & 7) * (G8 >> 3) + (b >> 3);
}

// NR right-hand sides that share the operator (indices o, o+nop, o+2 nop,
// ... of a batch) ride through one thread: the operator-only quantities
// (bands, pivots, the a~, c~ coefficients of the sweeps and the A, C of the
// reduced rows) are computed once for all of them, only the right-hand-side
// recurrences are repeated.  Block (mode tile, y): operator o = y % nop,
// right-hand sides o + nop (NR (y / nop) + r), r < NR, those past the batch
// are computed on a clamped index and not stored.
template <typename T, int MMAX, int NR, int NT, class Prob>
__global__ __launch_bounds__(NT) void k_tri(Prob prob, long nmodes, int W, int P, int batch, int nop) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // two buffers of the reduced rows, each [A | C | D_0 .. D_{NR-1}], rows x W
    C2<T> *lds = reinterpret_cast<C2<T> *>(smem_raw);
    const int RR = 2 * P;
    const int RW = RR * W;
    const int BUF = (2 + NR) * RW;
    const int n = prob.n();
    const int w = threadIdx.x % W, c = threadIdx.x / W;
    const long m = xcd_tile(blockIdx.x, gridDim.x) * W + w;
    const bool valid = (c < P) && (m < nmodes);
    const int o = blockIdx.y % nop, jg = blockIdx.y / nop;
    int bidx[NR];
    bool bok[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int b = o + nop * (jg * NR + r);
        bok[r] = b < batch;
        bidx[r] = bok[r] ? b : (o < batch ? o : 0);
    }
    // Idle lanes (c >= P, padding modes) run the same code on a clamped
    // chunk and mode so that every load is in bounds and no branch splits
    // the sweeps; their results are discarded.
    const int cq = c < P ? c : P - 1;
    const int s = (int)(((long)cq * n) / P), e = (int)(((long)(cq + 1) * n) / P) - 1;
    const int len = e - s + 1;   // 2 <= len <= MMAX
    const auto md = prob.mode(m < nmodes ? m : nmodes - 1, bidx[0]);

    // 0. issue every right-hand-side load of the chunk up front (rows past
    //    e repeat row e), so that the HBM latency is paid once per chunk and
    //    not once per row of the recurrence
    C2<T> db[NR][MMAX], ab[MMAX], cb[MMAX];
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
        for (int t = 0; t < MMAX; t++) db[r][t] = prob.rhs(bidx[r], md, min(s + t, e));
    const Band<T> bs = prob.band(md, s);
    const C2<T> sub_s = bs.sub, dia_s = bs.dia, sup_s = bs.sup;
    C2<T> d_s[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) d_s[r] = db[r][0];

    // 1. downward sweep over rows s+1..e: row s+t becomes
    //    ab[t] x_s + x_{s+t} + cb[t] x_{s+t+1} = db[t]
    C2<T> pa = mk<T>(0, 0), pc = mk<T>(0, 0), pd[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) pd[r] = mk<T>(0, 0);
#pragma unroll
    for (int t = 1; t < MMAX; t++) {
        const Band<T> bq = prob.band(md, min(s + t, e));
        const C2<T> a = bq.sub;
        const C2<T> den = (t == 1) ? bq.dia : cfms<T>(a, pc, bq.dia);
        const C2<T> rc = crcp<T>(den);
        const C2<T> na = (t == 1) ? cmul<T>(a, rc) : cneg<T>(cmul<T>(cmul<T>(a, pa), rc));
        const C2<T> nc = cmul<T>(bq.sup, rc);
        C2<T> nd[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) nd[r] = (t == 1) ? cmul<T>(db[r][t], rc) : cmul<T>(cfms<T>(a, pd[r], db[r][t]), rc);
        if (t < len) {   // selects, not a branch
            ab[t] = na; cb[t] = nc;
            pa = na; pc = nc;
#pragma unroll
            for (int r = 0; r < NR; r++) { db[r][t] = nd[r]; pd[r] = nd[r]; }
        }
    }
    // reduced row 2c+1 (x_e): pa x_s + x_e + pc x_{e+1} = pd  (already normalised)
    C2<T> A1 = pa, C1 = pc, D1[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) D1[r] = pd[r];
    // 2. upward sweep over rows e-2..s+1 (row e-1 keeps its downward form):
    //    row s+t becomes ab[t] x_s + x_{s+t} + cb[t] x_e = db[t]
#pragma unroll
    for (int t = MMAX - 3; t >= 1; t--) {
        const C2<T> cbt = cb[t];
        const C2<T> na = cfms<T>(cbt, ab[t + 1], ab[t]);
        const C2<T> nc = cneg<T>(cmul<T>(cbt, cb[t + 1]));
        C2<T> nd[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) nd[r] = cfms<T>(cbt, db[r][t + 1], db[r][t]);
        if (t <= len - 3) {
            ab[t] = na; cb[t] = nc;
#pragma unroll
            for (int r = 0; r < NR; r++) db[r][t] = nd[r];
        }
    }
    // reduced row 2c (x_s): sub_s x_{s-1} + B x_s + C x_e = D
    C2<T> B0, C0, D0[NR];
    if (len >= 3) {
        B0 = cfms<T>(sup_s, ab[1], dia_s);
        C0 = cneg<T>(cmul<T>(sup_s, cb[1]));
#pragma unroll
        for (int r = 0; r < NR; r++) D0[r] = cfms<T>(sup_s, db[r][1], d_s[r]);
    } else {
        B0 = dia_s; C0 = sup_s;
#pragma unroll
        for (int r = 0; r < NR; r++) D0[r] = d_s[r];
    }
    C2<T> A0;
    {
        const C2<T> rc = crcp<T>(B0);
        A0 = cmul<T>(sub_s, rc);
        C0 = cmul<T>(C0, rc);
#pragma unroll
        for (int r = 0; r < NR; r++) D0[r] = cmul<T>(D0[r], rc);
    }
    if (!valid) {  // identity rows for idle threads and padding modes
        A0 = C0 = A1 = C1 = mk<T>(0, 0);
#pragma unroll
        for (int r = 0; r < NR; r++) D0[r] = D1[r] = mk<T>(0, 0);
    }
    // 3. PCR on the reduced system, rows r0 = 2c, r1 = 2c+1, layout [row][W].
    //    Two LDS buffers: step i reads buffer i%2 and writes the other, so
    //    one barrier per step suffices.
    // state of the thread's two reduced rows q = 2c + h, h = 0 (x_s), 1 (x_e),
    // row q at LDS slot q W + w.  (Slots (q & 1) P + (q >> 1), free of the
    // two-way bank conflicts of this layout, measured 7% slower: 872 against
    // 811 us at N=1023 x 64 right-hand sides.)
    auto slot = [&](int q) { return q * W + w; };
    C2<T> PA[2] = {A0, A1}, PC[2] = {C0, C1}, PD[2][NR];
#pragma unroll
    for (int r = 0; r < NR; r++) { PD[0][r] = D0[r]; PD[1][r] = D1[r]; }
    if (c < P) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int q = 2 * c + h;
            lds[slot(q)] = PA[h];
            lds[RW + slot(q)] = PC[h];
#pragma unroll
            for (int r = 0; r < NR; r++) lds[(2 + r) * RW + slot(q)] = PD[h][r];
        }
    }
    int cur = 0;
    for (int dd = 1; dd < RR; dd <<= 1) {
        __syncthreads();
        const C2<T> *sb = lds + cur * BUF;
        C2<T> *tb = lds + (cur ^ 1) * BUF;
        if (c < P) {
            // row q: A x_{q-d} + x_q + C x_{q+d} = D, eliminated against rows q-d and q+d
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int q = 2 * c + h;
                // neighbours outside the system read the row itself and are zeroed
                const C2<T> z = mk<T>(0, 0);
                const int qL = q - dd, qU = q + dd;
                const bool hL = qL >= 0, hU = qU < RR;
                const int iL = slot(hL ? qL : q), iU = slot(hU ? qU : q);
                C2<T> aL = sb[iL], cL = sb[RW + iL], aU = sb[iU], cU = sb[RW + iU], dL[NR], dU[NR];
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    dL[r] = sb[(2 + r) * RW + iL];
                    dU[r] = sb[(2 + r) * RW + iU];
                }
                if (!hL) {
                    aL = z; cL = z;
#pragma unroll
                    for (int r = 0; r < NR; r++) dL[r] = z;
                }
                if (!hU) {
                    aU = z; cU = z;
#pragma unroll
                    for (int r = 0; r < NR; r++) dU[r] = z;
                }
                const C2<T> A = PA[h], C = PC[h];
                const C2<T> bb = csub<T>(mk<T>(1, 0), cadd<T>(cmul<T>(A, cL), cmul<T>(C, aU)));
                const C2<T> ib = crcp<T>(bb);
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    const C2<T> nd = csub<T>(PD[h][r], cadd<T>(cmul<T>(A, dL[r]), cmul<T>(C, dU[r])));
                    PD[h][r] = cmul<T>(nd, ib);
                    tb[(2 + r) * RW + slot(q)] = PD[h][r];
                }
                PA[h] = cmul<T>(cneg<T>(cmul<T>(A, aL)), ib);
                PC[h] = cmul<T>(cneg<T>(cmul<T>(C, cU)), ib);
                tb[slot(q)] = PA[h];
                tb[RW + slot(q)] = PC[h];
            }
        }
        cur ^= 1;
    }
#pragma unroll
    for (int r = 0; r < NR; r++) { D0[r] = PD[0][r]; D1[r] = PD[1][r]; }
    if (!valid) return;
    // 4. back-substitution of the interior and store
#pragma unroll
    for (int r = 0; r < NR; r++) {
        if (!bok[r]) continue;
        const C2<T> xs = D0[r], xe = D1[r];
        const int b = bidx[r];
        prob.store(b, md, s, xs);
#pragma unroll
        for (int t = 1; t < MMAX - 1; t++) {
            if (t < len - 1) {
                C2<T> x = csub<T>(db[r][t], cadd<T>(cmul<T>(ab[t], xs), cmul<T>(cb[t], xe)));
                prob.store(b, md, s + t, x);
            }
        }
        prob.store(b, md, e, xe);
    }
}

// LDS bytes of the reduced system: two buffers of A, C and NR right-hand sides.
inline size_t tri_lds_bytes(size_t elem, int P, int W, int nr) { return elem * 2 * (2 + nr) * 2 * (size_t)P * W; }

template <typename T, int MM, int NR, class Prob>
static hipError_t launch_tri_nr(const Prob &pb, int n, long nmodes, int batch, int nop, int W, int P,
                                hipStream_t st) {
    if (n < 2 || batch < 1 || nop < 1) return hipErrorInvalidValue;
    // every chunk needs at least two rows, at most MM rows
    if (P < 1 || P > n / 2 || (n + P - 1) / P > MM || W * P > 1024) return hipErrorInvalidValue;
    int threads = ((W * P + 63) / 64) * 64;
    const size_t shm = tri_lds_bytes(sizeof(C2<T>), P, W, NR);
    if (shm > 160 * 1024) return hipErrorInvalidValue;
    // 512-thread bound: at most 256 VGPRs per lane and two waves per SIMD.
    // (16 FP64 rows spill a few registers there; measured faster than the
    // spill-free 256-thread bound, which leaves one wave per SIMD.)
    auto kfn = threads <= 512 ? k_tri<T, MM, NR, 512, Prob> : k_tri<T, MM, NR, 1024, Prob>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    const int per_op = (batch + nop - 1) / nop;
    const int groups = (per_op + NR - 1) / NR;
    dim3 grid((unsigned)((nmodes + W - 1) / W), (unsigned)(nop * groups));
    hipLaunchKernelGGL(kfn, grid, dim3(threads), shm, st, pb, nmodes, W, P, batch, nop);
    return hipGetLastError();
}

template <typename T, int MM, class Prob>
static hipError_t launch_tri_prob(const Prob &pb, int n, long nmodes, int batch, int nop, int W, int P, int nr,
                                  hipStream_t st) {
    if (nr == 1) return launch_tri_nr<T, MM, 1>(pb, n, nmodes, batch, nop, W, P, st);
    if constexpr (sizeof(T) == 8) {   // several right-hand sides per thread: FP64 elimination only
        if (nr == 2) return launch_tri_nr<T, MM, 2>(pb, n, nmodes, batch, nop, W, P, st);
        if (nr == 4) return launch_tri_nr<T, MM, 4>(pb, n, nmodes, batch, nop, W, P, st);
    }
    return hipErrorInvalidValue;
}

void tri_shape(int n, long nmodes, int prec, int *W, int *P, int *mmax) {
    // Measured on MI355X (FP64, k = 50 compact operators, row-major field):
    // 256-thread blocks win at every length tried; 16 rows per thread above
    // n = 512 (fewer chunks, a shallower reduced system), 8 below.  n = 1023,
    // 64 x 1023 modes: W = 4 x 16 rows 809 us against 899 us for W = 2 x 8
    // rows and 917 us for W = 8 x 16; n = 255, 255^2 modes: W = 8 x 8 rows
    // 207 us against 283 us for 16 x 8.
    int mm = n > 512 ? 16 : 8;
    if (n <= 16) mm = 4;
    (void)prec;
    int p = (n + mm - 1) / mm;
    if (p > n / 2) p = n / 2;
    if (p < 1) p = 1;
    int w = 256 / p;
    if (w > 16) w = 16;
    if (w < 1) w = 1;
    int pw = 1;
    while (pw * 2 <= w) pw *= 2;
    w = pw;
    while (w > 1 && w / 2 >= nmodes) w /= 2;
    *W = w; *P = p; *mmax = mm;
}

template <typename T, typename Ti, typename To, int MM, int DIM>
static hipError_t launch_tri_d(const TriHelm &t, const void *in, void *out, hipStream_t st) {
    HelmProb<T, Ti, To, DIM> pb;
    pb.bands.Nx = t.Nx;
    pb.bands.coef = (const C2<T> *)t.coef;
    pb.bands.nop = t.nop > 0 ? t.nop : 1;
    pb.bands.cstride = t.cstride;
    pb.bands.lamx = t.lamx;
    pb.bands.lamy = t.lamy;
    pb.in = (const C2<Ti> *)in;
    pb.out = (C2<To> *)out;
    pb.field = t.field;
    long nmodes;
    if (t.dim == 2) { pb.bands.n = t.Ny; nmodes = t.Nx; pb.S = t.Nx; }
    else { pb.bands.n = t.Nz; nmodes = (long)t.Nx * t.Ny; pb.S = (long)t.Nx * t.Ny; }
    pb.nmodes = nmodes;
    // tiles are powers of two holding whole blocks of W modes
    pb.tw = t.tiled;
    if (t.tiled && ((t.tiled & (t.tiled - 1)) || t.tiled % t.W)) return hipErrorInvalidValue;
    return launch_tri_prob<T, MM>(pb, pb.bands.n, nmodes, t.batch, pb.bands.nop, t.W, t.P, t.nr, st);
}

template <typename T, typename Ti, typename To, int MM>
static hipError_t launch_tri_t(const TriHelm &t, const void *in, void *out, hipStream_t st) {
    if (t.dim == 2) return launch_tri_d<T, Ti, To, MM, 2>(t, in, out, st);
    return launch_tri_d<T, Ti, To, MM, 3>(t, in, out, st);
}

template <typename T, typename Ti, typename To>
static hipError_t launch_tri_mm(const TriHelm &t, const void *in, void *out, hipStream_t st) {
    switch (t.mmax) {
    case 4: return launch_tri_t<T, Ti, To, 4>(t, in, out, st);
    case 8: return launch_tri_t<T, Ti, To, 8>(t, in, out, st);
    case 16: return launch_tri_t<T, Ti, To, 16>(t, in, out, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_compact(const Compact &cd, hipStream_t st) {
    CompactProb<double> pb;
    pb.nn = cd.n;
    pb.first = cd.first;
    pb.diag = cd.diag;
    pb.h = cd.h;
    pb.f = (const double2 *)cd.f;
    pb.field_in = cd.field_in;
    pb.field_out = cd.field_out;
    pb.field_bnd = cd.field_bnd;
    pb.mA = cd.mA; pb.orig = cd.orig; pb.sA = cd.sA; pb.sB = cd.sB; pb.eS = cd.eS;
    pb.lo = (const double2 *)cd.lo; pb.hi = (const double2 *)cd.hi;
    pb.bo = cd.bo; pb.bA = cd.bA; pb.bB = cd.bB;
    pb.out = (double2 *)cd.out;
    pb.oorig = cd.oorig; pb.oA = cd.oA; pb.oB = cd.oB; pb.oS = cd.oS;
    int W, P, mm;
    tri_shape(cd.n, cd.nmodes, F64, &W, &P, &mm);
    switch (mm) {
    case 4: return launch_tri_nr<double, 4, 1>(pb, cd.n, cd.nmodes, cd.batch, 1, W, P, st);
    case 8: return launch_tri_nr<double, 8, 1>(pb, cd.n, cd.nmodes, cd.batch, 1, W, P, st);
    case 16: return launch_tri_nr<double, 16, 1>(pb, cd.n, cd.nmodes, cd.batch, 1, W, P, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_tri_helm(const TriHelm &t, DstTypes ty, const void *in, void *out, hipStream_t st) {
    if (ty.comp == 0 && ty.in == 0 && ty.out == 0) return launch_tri_mm<double, double, double>(t, in, out, st);
    if (ty.comp == 1 && ty.in == 1 && ty.out == 1) return launch_tri_mm<float, float, float>(t, in, out, st);
    if (ty.comp == 0 && ty.in == 1 && ty.out == 1) return launch_tri_mm<double, float, float>(t, in, out, st);
    return hipErrorInvalidValue;
}

}  // namespace hh
