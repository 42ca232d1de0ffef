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
#include <type_traits>

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
template <typename OT> struct Band {
    OT sub, dia, sup;
};

// Arithmetic on the operator's coefficients, OT = C2<T> (complex bands) or
// T (real bands: every stencil coefficient real, as for real k), and on
// complex right-hand sides scaled by them.  The real form needs a quarter
// of the multiplies of the complex one and half the registers and LDS.
template <typename T, typename OT> struct OpAr;
template <typename T> struct OpAr<T, C2<T>> {
    using O = C2<T>;
    static HH_DEV O zero() { return mk<T>(0, 0); }
    static HH_DEV O one() { return mk<T>(1, 0); }
    static HH_DEV O mul(O a, O b) { return cmul<T>(a, b); }
    static HH_DEV O neg(O a) { return cneg<T>(a); }
    static HH_DEV O fms(O a, O b, O c) { return cfms<T>(a, b, c); }      // c - a b
    static HH_DEV O rcp(O a) { return crcp<T>(a); }
    static HH_DEV C2<T> xmul(O a, C2<T> x) { return cmul<T>(a, x); }
    static HH_DEV C2<T> xfms(O a, C2<T> x, C2<T> y) { return cfms<T>(a, x, y); }   // y - a x
    // y - (a x + b z), c - (a b + d e): the products summed first
    static HH_DEV C2<T> xfms2(O a, C2<T> x, O b, C2<T> z, C2<T> y) {
        return csub<T>(y, cadd<T>(cmul<T>(a, x), cmul<T>(b, z)));
    }
    static HH_DEV O fms2(O a, O b, O d, O e, O c) { return csub<T>(c, cadd<T>(cmul<T>(a, b), cmul<T>(d, e))); }
};
template <typename T> struct OpAr<T, T> {
    using O = T;
    static HH_DEV O zero() { return (T)0; }
    static HH_DEV O one() { return (T)1; }
    static HH_DEV O mul(O a, O b) { return a * b; }
    static HH_DEV O neg(O a) { return -a; }
    static HH_DEV O fms(O a, O b, O c) { return c - a * b; }
    // rcp_nr's Newton steps hold for either sign of a nonzero argument
    static HH_DEV O rcp(O a) { return rcp_nr(a); }
    static HH_DEV C2<T> xmul(O a, C2<T> x) { return mk<T>(a * x.x, a * x.y); }
    static HH_DEV C2<T> xfms(O a, C2<T> x, C2<T> y) { return mk<T>(y.x - a * x.x, y.y - a * x.y); }
    static HH_DEV C2<T> xfms2(O a, C2<T> x, O b, C2<T> z, C2<T> y) { return xfms(b, z, xfms(a, x, y)); }
    static HH_DEV O fms2(O a, O b, O d, O e, O c) { return fms(d, e, fms(a, b, c)); }
};

template <typename T, int DIM, bool REAL> struct HelmBands {
    using OT = typename std::conditional<REAL, T, C2<T>>::type;
    int n, Nx, nop;
    // complex table [q][6] / [q][12]; real table: the same coefficients'
    // real parts in pairs, [q][3] / [q][6] of (re c_{2i}, re c_{2i+1})
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
    // Bands in the reference's operation order with no fused multiply-adds
    // (2D_Program/main_functions/LU.c:16-23, 3D_Program/main_functions/
    // LU.c:34-47): the FP64 bands equal the reference's to the last bit.
    HH_DEV Band<OT> band(const Mode &md, int q) const {
#pragma clang fp contract(off)
        Band<OT> r;
        const T ms = q > 0 ? (T)1 : (T)0, ps = q < n - 1 ? (T)1 : (T)0;
        if constexpr (REAL) {
            if constexpr (DIM == 2) {   // pairs (bm, am), (b, a), (bp, ap)
                const C2<T> *c = md.cf + 3 * (long)q;
                const T l = md.l0;
                const C2<T> c0 = c[0], c1 = c[1], c2 = c[2];
                r.sub = ms * (c0.x * l + c0.y);
                r.dia = c1.x * l + c1.y;
                r.sup = ps * (c2.x * l + c2.y);
            } else {   // pairs (a, b), (c, d), (ap, bp), (cp, dp), (am, bm), (cm, dm)
                const C2<T> *c = md.cf + 6 * (long)q;
                const T a = md.l0, b = md.l1, g = md.l2;
                const C2<T> c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3], c4 = c[4], c5 = c[5];
                r.dia = c0.x * a + c0.y * b + c1.x * g + c1.y;
                r.sup = ps * (c2.x * a + c2.y * b + c3.x * g + c3.y);
                r.sub = ms * (c4.x * a + c4.y * b + c5.x * g + c5.y);
            }
        } else if constexpr (DIM == 2) {
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
template <typename T, typename Ti, typename To, int DIM, bool REAL> struct HelmProb {
    using OT = typename HelmBands<T, DIM, REAL>::OT;
    HelmBands<T, DIM, REAL> bands;
    const C2<Ti> *in;
    C2<To> *out;
    long field, S, nmodes;
    int tw;
    const uint32_t *fbits;   // near-resonant modes, solved by k_tri_seq: not stored here
    int fwords;
    struct Mode { typename HelmBands<T, DIM, REAL>::Mode bm; long off, step; };
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
    HH_DEV Band<OT> band(const Mode &md, int q) const { return bands.band(md.bm, q); }
    HH_DEV C2<T> rhs(int b, const Mode &md, int q) const {
        return cconv<T, Ti>(in[b * field + md.off + (long)q * md.step]);
    }
    HH_DEV void store(int b, const Mode &md, int q, C2<T> x) const {
        out[b * field + md.off + (long)q * md.step] = cconv<To, T>(x);
    }
    HH_DEV bool skip(int o, long m) const { return fbits && ((fbits[(long)o * fwords + (m >> 5)] >> (m & 31)) & 1u); }
    // FP64 access of the sequential kernel
    HH_DEV double2 rhs64(int b, const Mode &md, int q) const {
        return cconv<double, Ti>(in[b * field + md.off + (long)q * md.step]);
    }
    HH_DEV void store64(int b, const Mode &md, int q, double2 x) const {
        out[b * field + md.off + (long)q * md.step] = cconv<To, double>(x);
    }
    // one part (comp 0 real, 1 imaginary) of an element, as FP64
    HH_DEV double rhs64c(int b, const Mode &md, int q, int comp) const {
        return (double)reinterpret_cast<const Ti *>(in + b * field + md.off + (long)q * md.step)[comp];
    }
    HH_DEV void store64c(int b, const Mode &md, int q, int comp, double v) const {
        reinterpret_cast<To *>(out + b * field + md.off + (long)q * md.step)[comp] = (To)v;
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
    using OT = T;   // bands (1, diag, 1)
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
        r.sub = q > 0 ? (T)1 : (T)0;
        r.dia = diag;
        r.sup = q < nn - 1 ? (T)1 : (T)0;
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
    HH_DEV bool skip(int, long) const { return false; }
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
__global__ __launch_bounds__(NT) void k_tri(Prob prob, long nmodes, int W, int P, int batch, int nop, int g0) {
    using OT = typename Prob::OT;
    using O = OpAr<T, OT>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // two buffers of the reduced rows, each [A | C | D_0 .. D_{NR-1}], rows x W;
    // A and C are operator values (OT), the D the right-hand sides (complex)
    const int RR = 2 * P;
    const int RW = RR * W;
    const int BUFB = RW * (int)(2 * sizeof(OT) + NR * sizeof(C2<T>));   // bytes per buffer (< 160 KiB)
    auto bufA = [&](int k) { return reinterpret_cast<OT *>(smem_raw + k * BUFB); };
    auto bufC = [&](int k) { return reinterpret_cast<OT *>(smem_raw + k * BUFB) + RW; };
    auto bufD = [&](int k) { return reinterpret_cast<C2<T> *>(reinterpret_cast<OT *>(smem_raw + k * BUFB) + 2 * RW); };
    const int n = prob.n();
    const int w = threadIdx.x % W, c = threadIdx.x / W;
    const long m = xcd_tile(blockIdx.x, gridDim.x) * W + w;
    const bool valid = (c < P) && (m < nmodes);
    // right-hand-side groups in reverse launch order: the row transform
    // before this pass wrote the last right-hand sides last, so they are the
    // ones still in the memory-side cache when the pass starts (and the
    // first ones this pass writes last, where the next row transform starts)
    const int ng = (int)(gridDim.y / nop);
    const int o = blockIdx.y % nop, jg = g0 + (ng - 1 - (int)(blockIdx.y / nop));
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
    C2<T> db[NR][MMAX];
    OT ab[MMAX], cb[MMAX];
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
        for (int t = 0; t < MMAX; t++) db[r][t] = prob.rhs(bidx[r], md, min(s + t, e));
    const Band<OT> bs = prob.band(md, s);
    const OT sub_s = bs.sub, dia_s = bs.dia, sup_s = bs.sup;
    C2<T> d_s[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) d_s[r] = db[r][0];

    // 1. downward sweep over rows s+1..e: row s+t becomes
    //    ab[t] x_s + x_{s+t} + cb[t] x_{s+t+1} = db[t]
    OT pa = O::zero(), pc = O::zero();
    C2<T> pd[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) pd[r] = mk<T>(0, 0);
#pragma unroll
    for (int t = 1; t < MMAX; t++) {
        const Band<OT> bq = prob.band(md, min(s + t, e));
        const OT a = bq.sub;
        const OT den = (t == 1) ? bq.dia : O::fms(a, pc, bq.dia);
        const OT rc = O::rcp(den);
        const OT na = (t == 1) ? O::mul(a, rc) : O::neg(O::mul(O::mul(a, pa), rc));
        const OT nc = O::mul(bq.sup, rc);
        C2<T> nd[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) nd[r] = (t == 1) ? O::xmul(rc, db[r][t]) : O::xmul(rc, O::xfms(a, pd[r], db[r][t]));
        if (t < len) {   // selects, not a branch
            ab[t] = na; cb[t] = nc;
            pa = na; pc = nc;
#pragma unroll
            for (int r = 0; r < NR; r++) { db[r][t] = nd[r]; pd[r] = nd[r]; }
        }
    }
    // reduced row 2c+1 (x_e): pa x_s + x_e + pc x_{e+1} = pd  (already normalised)
    OT A1 = pa, C1 = pc;
    C2<T> D1[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) D1[r] = pd[r];
    // 2. upward sweep over rows e-2..s+1 (row e-1 keeps its downward form):
    //    row s+t becomes ab[t] x_s + x_{s+t} + cb[t] x_e = db[t]
#pragma unroll
    for (int t = MMAX - 3; t >= 1; t--) {
        const OT cbt = cb[t];
        const OT na = O::fms(cbt, ab[t + 1], ab[t]);
        const OT nc = O::neg(O::mul(cbt, cb[t + 1]));
        C2<T> nd[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) nd[r] = O::xfms(cbt, db[r][t + 1], db[r][t]);
        if (t <= len - 3) {
            ab[t] = na; cb[t] = nc;
#pragma unroll
            for (int r = 0; r < NR; r++) db[r][t] = nd[r];
        }
    }
    // reduced row 2c (x_s): sub_s x_{s-1} + B x_s + C x_e = D
    OT B0, C0;
    C2<T> D0[NR];
    if (len >= 3) {
        B0 = O::fms(sup_s, ab[1], dia_s);
        C0 = O::neg(O::mul(sup_s, cb[1]));
#pragma unroll
        for (int r = 0; r < NR; r++) D0[r] = O::xfms(sup_s, db[r][1], d_s[r]);
    } else {
        B0 = dia_s; C0 = sup_s;
#pragma unroll
        for (int r = 0; r < NR; r++) D0[r] = d_s[r];
    }
    OT A0;
    {
        const OT rc = O::rcp(B0);
        A0 = O::mul(sub_s, rc);
        C0 = O::mul(C0, rc);
#pragma unroll
        for (int r = 0; r < NR; r++) D0[r] = O::xmul(rc, D0[r]);
    }
    if (!valid) {  // identity rows for idle threads and padding modes
        A0 = C0 = A1 = C1 = O::zero();
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
    OT PA[2] = {A0, A1}, PC[2] = {C0, C1};
    C2<T> PD[2][NR];
#pragma unroll
    for (int r = 0; r < NR; r++) { PD[0][r] = D0[r]; PD[1][r] = D1[r]; }
    if (c < P) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int q = 2 * c + h;
            bufA(0)[slot(q)] = PA[h];
            bufC(0)[slot(q)] = PC[h];
#pragma unroll
            for (int r = 0; r < NR; r++) bufD(0)[r * RW + slot(q)] = PD[h][r];
        }
    }
    int cur = 0;
    for (int dd = 1; dd < RR; dd <<= 1) {
        __syncthreads();
        const OT *sA = bufA(cur), *sC = bufC(cur);
        const C2<T> *sD = bufD(cur);
        OT *tA = bufA(cur ^ 1), *tC = bufC(cur ^ 1);
        C2<T> *tD = bufD(cur ^ 1);
        if (c < P) {
            // row q: A x_{q-d} + x_q + C x_{q+d} = D, eliminated against rows q-d and q+d
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int q = 2 * c + h;
                // neighbours outside the system read the row itself and are zeroed
                const int qL = q - dd, qU = q + dd;
                const bool hL = qL >= 0, hU = qU < RR;
                const int iL = slot(hL ? qL : q), iU = slot(hU ? qU : q);
                OT aL = sA[iL], cL = sC[iL], aU = sA[iU], cU = sC[iU];
                C2<T> dL[NR], dU[NR];
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    dL[r] = sD[r * RW + iL];
                    dU[r] = sD[r * RW + iU];
                }
                if (!hL) {
                    aL = cL = O::zero();
#pragma unroll
                    for (int r = 0; r < NR; r++) dL[r] = mk<T>(0, 0);
                }
                if (!hU) {
                    aU = cU = O::zero();
#pragma unroll
                    for (int r = 0; r < NR; r++) dU[r] = mk<T>(0, 0);
                }
                const OT A = PA[h], C = PC[h];
                const OT bb = O::fms2(A, cL, C, aU, O::one());
                const OT ib = O::rcp(bb);
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    PD[h][r] = O::xmul(ib, O::xfms2(A, dL[r], C, dU[r], PD[h][r]));
                    tD[r * RW + slot(q)] = PD[h][r];
                }
                PA[h] = O::neg(O::mul(O::mul(A, aL), ib));
                PC[h] = O::neg(O::mul(O::mul(C, cU), ib));
                tA[slot(q)] = PA[h];
                tC[slot(q)] = PC[h];
            }
        }
        cur ^= 1;
    }
#pragma unroll
    for (int r = 0; r < NR; r++) { D0[r] = PD[0][r]; D1[r] = PD[1][r]; }
    if (!valid || prob.skip(o, m)) return;
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
                const C2<T> x = O::xfms2(ab[t], xs, cb[t], xe, db[r][t]);
                prob.store(b, md, s + t, x);
            }
        }
        prob.store(b, md, e, xe);
    }
}

// Near-resonant modes: the reference's substitution (2D_Program/
// solver_functions/solver.c:71-86, 3D_Program/solver_functions/
// solver.c:356-369) in its own operation order, with the factors the host
// computed in the order of its LU():
//   y_0 = d_0,  y_q = d_q - L_q y_{q-1};
//   2D: x_q = (y_q - Up_q x_{q+1}) / U_q;   3D: x_q = (y_q - Up_q x_{q+1}) U_q,
// FP64 with no fused operations, whatever the storage type, so that for
// real operators every operation rounds as the reference's does and the
// result differs from the reference's only through its input (the
// transforms' rounding).  The 2D quotient is the correctly rounded one:
// q0 = a r, q = q0 + (a - p q0) r with r = RN(1/p) from the host and the
// residual exact by a fused multiply-add (Markstein).
//
// The recurrences are sequential, so the parallelism is across chains (one
// chain = one near-resonant mode of one right-hand side): a block runs C
// chains side by side, lane l holding chain l % C and, at position l / C
// of its wave, R consecutive rows in registers (all loads issued up
// front).  The chains' positions are walked in order: the C lanes of one
// position run their rows while the others wait, then hand the carried
// value to the next position by a lane shuffle, and from wave to wave
// through LDS.  Block x runs items x C .. x C + C - 1, item = b nfx + e_o,
// right-hand side b, its operator's near-resonant entry e_o.
template <typename OT> struct SeqOps;
template <> struct SeqOps<double> {   // real factors, complex right-hand sides
    static HH_DEV double zero() { return 0.0; }
    static HH_DEV double one() { return 1.0; }
    static HH_DEV double2 fms(double a, double2 x, double2 y) {   // y - a x
#pragma clang fp contract(off)
        return mk<double>(y.x - a * x.x, y.y - a * x.y);
    }
    static HH_DEV double2 mul(double a, double2 x) { return mk<double>(a * x.x, a * x.y); }
    static HH_DEV double q1(double a, double p, double r) {
        const double q0 = a * r;
        return __builtin_fma(__builtin_fma(-p, q0, a), r, q0);
    }
    static HH_DEV double2 div(double2 a, double p, double r) { return mk<double>(q1(a.x, p, r), q1(a.y, p, r)); }
};
template <> struct SeqOps<double2> {   // complex factors: C99 products, quotients through the reciprocal
    static HH_DEV double2 zero() { return mk<double>(0, 0); }
    static HH_DEV double2 one() { return mk<double>(1, 0); }
    static HH_DEV double2 mul(double2 a, double2 x) {
#pragma clang fp contract(off)
        return mk<double>(a.x * x.x - a.y * x.y, a.x * x.y + a.y * x.x);
    }
    static HH_DEV double2 fms(double2 a, double2 x, double2 y) {
#pragma clang fp contract(off)
        const double2 p = mul(a, x);
        return mk<double>(y.x - p.x, y.y - p.y);
    }
    static HH_DEV double2 div(double2 a, double2, double2 r) { return mul(r, a); }
};

HH_DEV double2 shfl_up2(double2 v, int k) { return mk<double>(__shfl_up(v.x, k, 64), __shfl_up(v.y, k, 64)); }
HH_DEV double2 shfl_down2(double2 v, int k) { return mk<double>(__shfl_down(v.x, k, 64), __shfl_down(v.y, k, 64)); }

template <int R, int C, bool DIV, class Prob, typename OT>
__global__ __launch_bounds__(512) void k_tri_seq(Prob prob, const int *foff, const int *fmode, const OT *fac,
                                                 long nent, int nfx, int batch, int nop) {
    using S = SeqOps<OT>;
    constexpr int PW = 64 / C;   // positions per wave
    __shared__ double2 carry[C];
    const int n = prob.n();
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, NW = blockDim.x >> 6;
    const int c = lane % C, pw = lane / C;
    const int s = (wave * PW + pw) * R;   // first row of this lane
    const long item = (long)blockIdx.x * C + c;
    const int b = (int)(item / nfx), x = (int)(item % nfx);
    const int o = b < batch ? b % nop : 0;
    const bool live = b < batch && x < foff[o + 1] - foff[o];
    // idle chains of the last block run on a valid item and store nothing
    const int bb = live ? b : 0, e = live ? foff[o] + x : 0;
    const auto md = prob.mode(fmode[e], bb);
    // Rows past n (the tail of the last position) become identity rows
    // (d = 0, L = 0, P = R = 1, Up = 0): the forward walk ends on them with a
    // carry nobody reads, the backward walk leaves them with 0, as it starts,
    // so every real row keeps its bits, and the walks need no per-row branch.
    double2 d[R];
    OT L[R], Pv[R], Rv[R], Up[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
        const int q = min(s + i, n - 1);
        const bool pad = s + i >= n;
        const double2 v = prob.rhs64(bb, md, q);
        const OT *f = fac + ((long)q * nent + e) * 4;
        const OT f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3];
        d[i] = pad ? mk<double>(0, 0) : v;
        L[i] = pad ? S::zero() : f0;
        Pv[i] = pad ? S::one() : f1;
        Rv[i] = pad ? S::one() : f2;
        Up[i] = pad ? S::zero() : f3;
    }
    // forward: y_q = d_q - L_q y_{q-1} (L_0 = 0)
    double2 cy = mk<double>(0, 0);
    for (int w = 0; w < NW; w++) {
        if (wave == w) {
            if (w > 0) cy = carry[c];
            for (int pp = 0; pp < PW; pp++) {
                if (pw == pp) {
#pragma unroll
                    for (int i = 0; i < R; i++) {
                        d[i] = S::fms(L[i], cy, d[i]);
                        cy = d[i];
                    }
                    if (pp == PW - 1) carry[c] = cy;
                }
                if (pp < PW - 1) cy = shfl_up2(cy, C);
            }
        }
        __syncthreads();
    }
    // backward: x_q = (y_q - Up_q x_{q+1}) / U_q (2D) or * U_q (3D) (Up_{n-1} = 0)
    cy = mk<double>(0, 0);
    for (int w = NW - 1; w >= 0; w--) {
        if (wave == w) {
            if (w < NW - 1) cy = carry[c];
            for (int pp = PW - 1; pp >= 0; pp--) {
                if (pw == pp) {
#pragma unroll
                    for (int i = R - 1; i >= 0; i--) {
                        const double2 a = S::fms(Up[i], cy, d[i]);
                        if constexpr (DIV) d[i] = S::div(a, Pv[i], Rv[i]);
                        else d[i] = S::mul(Rv[i], a);
                        cy = d[i];
                    }
                    if (pp == 0) carry[c] = cy;
                }
                if (pp > 0) cy = shfl_down2(cy, C);
            }
        }
        __syncthreads();
    }
    if (!live) return;
#pragma unroll
    for (int i = 0; i < R; i++)
        if (s + i < n) prob.store64(b, md, s + i, d[i]);
}

// Real factors: the real and imaginary parts of a right-hand side follow
// two independent real recurrences with the same factors, so neighbouring
// lanes run them as separate chains (chain = 2 item + part): half the
// instructions per row on the walk's critical path, each operation the one
// k_tri_seq applies to that part (same bits).
HH_DEV double fms1(double a, double x, double y) {   // y - a x, unfused
#pragma clang fp contract(off)
    return y - a * x;
}

template <int R, int C, bool DIV, class Prob>
__global__ __launch_bounds__(512) void k_tri_seq_re(Prob prob, const int *foff, const int *fmode, const double *fac,
                                                    long nent, int nfx, int batch, int nop) {
    constexpr int PW = 64 / C;   // positions per wave
    __shared__ double carry[C];
    const int n = prob.n();
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, NW = blockDim.x >> 6;
    const int c = lane % C, pw = lane / C;
    const int s = (wave * PW + pw) * R;   // first row of this lane
    const long chain = (long)blockIdx.x * C + c;
    const long item = chain >> 1;
    const int comp = (int)(chain & 1);
    const int b = (int)(item / nfx), x = (int)(item % nfx);
    const int o = b < batch ? b % nop : 0;
    const bool live = b < batch && x < foff[o + 1] - foff[o];
    const int bb = live ? b : 0, e = live ? foff[o] + x : 0;
    const auto md = prob.mode(fmode[e], bb);
    // rows past n: identity rows, as in k_tri_seq
    double d[R], L[R], Pv[R], Rv[R], Up[R];
#pragma unroll
    for (int i = 0; i < R; i++) {
        const int q = min(s + i, n - 1);
        const bool pad = s + i >= n;
        const double v = prob.rhs64c(bb, md, q, comp);
        const double *f = fac + ((long)q * nent + e) * 4;
        const double f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3];
        d[i] = pad ? 0.0 : v;
        L[i] = pad ? 0.0 : f0;
        Pv[i] = pad ? 1.0 : f1;
        Rv[i] = pad ? 1.0 : f2;
        Up[i] = pad ? 0.0 : f3;
    }
    double cy = 0.0;
    for (int w = 0; w < NW; w++) {
        if (wave == w) {
            if (w > 0) cy = carry[c];
            for (int pp = 0; pp < PW; pp++) {
                if (pw == pp) {
#pragma unroll
                    for (int i = 0; i < R; i++) {
                        d[i] = fms1(L[i], cy, d[i]);
                        cy = d[i];
                    }
                    if (pp == PW - 1) carry[c] = cy;
                }
                if (pp < PW - 1) cy = __shfl_up(cy, C, 64);
            }
        }
        __syncthreads();
    }
    cy = 0.0;
    for (int w = NW - 1; w >= 0; w--) {
        if (wave == w) {
            if (w < NW - 1) cy = carry[c];
            for (int pp = PW - 1; pp >= 0; pp--) {
                if (pw == pp) {
#pragma unroll
                    for (int i = R - 1; i >= 0; i--) {
                        const double a = fms1(Up[i], cy, d[i]);
                        if constexpr (DIV) d[i] = SeqOps<double>::q1(a, Pv[i], Rv[i]);
                        else d[i] = Rv[i] * a;
                        cy = d[i];
                    }
                    if (pp == 0) carry[c] = cy;
                }
                if (pp > 0) cy = __shfl_down(cy, C, 64);
            }
        }
        __syncthreads();
    }
    if (!live) return;
#pragma unroll
    for (int i = 0; i < R; i++)
        if (s + i < n) prob.store64c(b, md, s + i, comp, d[i]);
}

template <typename T, int MM, int NR, class Prob>
static hipError_t launch_tri_nr(const Prob &pb, int n, long nmodes, int batch, int nop, int W, int P,
                                hipStream_t st) {
    if (n < 2 || batch < 1 || nop < 1) return hipErrorInvalidValue;
    // every chunk needs at least two rows, at most MM rows
    if (P < 1 || P > n / 2 || (n + P - 1) / P > MM || W * P > 1024) return hipErrorInvalidValue;
    int threads = ((W * P + 63) / 64) * 64;
    const size_t shm = tri_lds_bytes(sizeof(typename Prob::OT), sizeof(C2<T>), P, W, NR);
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
    // grid y holds at most 65535 blocks: launch the right-hand-side groups in
    // slices, the last slice first (the kernel walks its groups backwards too)
    if (nop > 65535) return hipErrorInvalidValue;
    const int gmax = 65535 / nop;
    for (int g0 = ((groups - 1) / gmax) * gmax; g0 >= 0; g0 -= gmax) {
        const int ng = groups - g0 < gmax ? groups - g0 : gmax;
        dim3 grid((unsigned)((nmodes + W - 1) / W), (unsigned)(nop * ng));
        hipLaunchKernelGGL(kfn, grid, dim3(threads), shm, st, pb, nmodes, W, P, batch, nop, g0);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
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
    tri_shape_rows(n, nmodes, mm, W, P);
    *mmax = mm;
}

void tri_shape_rows(int n, long nmodes, int mm, int *W, int *P) {
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
    *W = w; *P = p;
}

template <typename T, typename Ti, typename To, int MM, int DIM, bool REAL>
static hipError_t launch_tri_r(const TriHelm &t, const void *in, void *out, hipStream_t st);

template <typename T, typename Ti, typename To, int MM, int DIM>
static hipError_t launch_tri_d(const TriHelm &t, const void *in, void *out, hipStream_t st) {
    if (t.real) return launch_tri_r<T, Ti, To, MM, DIM, true>(t, in, out, st);
    return launch_tri_r<T, Ti, To, MM, DIM, false>(t, in, out, st);
}

template <typename T, typename Ti, typename To, int MM, int DIM, bool REAL>
static hipError_t launch_tri_r(const TriHelm &t, const void *in, void *out, hipStream_t st) {
    HelmProb<T, Ti, To, DIM, REAL> pb;
    pb.bands.Nx = t.Nx;
    // the real table holds the coefficients' real parts in pairs: half the entries
    pb.bands.coef = (const C2<T> *)(REAL ? t.coefr : t.coef);
    pb.bands.nop = t.nop > 0 ? t.nop : 1;
    pb.bands.cstride = REAL ? t.cstride / 2 : t.cstride;
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
    pb.fbits = t.fbits;
    pb.fwords = t.fwords;
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

template <int R, int C, typename Ti, typename To, int DIM, bool REAL>
static hipError_t launch_seq_rc(const TriHelm &t, const void *in, void *out, hipStream_t st) {
    using OT = typename std::conditional<REAL, double, double2>::type;
    HelmProb<double, Ti, To, DIM, REAL> pb;
    pb.bands.Nx = t.Nx;
    pb.bands.coef = (const double2 *)(REAL ? t.coefr : t.coef);
    pb.bands.nop = t.nop > 0 ? t.nop : 1;
    pb.bands.cstride = REAL ? t.cstride / 2 : t.cstride;
    pb.bands.lamx = t.lamx;
    pb.bands.lamy = t.lamy;
    pb.in = (const C2<Ti> *)in;
    pb.out = (C2<To> *)out;
    pb.field = t.field;
    if (t.dim == 2) { pb.bands.n = t.Ny; pb.nmodes = t.Nx; pb.S = t.Nx; }
    else { pb.bands.n = t.Nz; pb.nmodes = (long)t.Nx * t.Ny; pb.S = (long)t.Nx * t.Ny; }
    pb.tw = t.tiled;
    pb.fbits = nullptr;
    pb.fwords = 0;
    const int n = pb.bands.n;
    constexpr int PW = 64 / C;
    const int G = ((n + R - 1) / R + PW - 1) / PW * PW;   // positions, whole waves
    const int NT = G * C;
    if (NT > 512) return hipErrorInvalidValue;
    const long items = (long)t.nfx * t.batch;
    // Real factors: one chain per part (k_tri_seq_re) when the doubled grid
    // still runs as one wave of blocks, one block per CU (a block holds a
    // CU's registers); otherwise one chain per item.  C2's 1152 items: the
    // split grid's 288 blocks ran in two rounds, 79.6 against 64 us.
    int ncu = 0;
    if constexpr (REAL) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            ncu = 0;
    }
    if (REAL && (2 * items + C - 1) / C <= ncu) {   // two chains per item: its real and imaginary parts
        const long blocks = (2 * items + C - 1) / C;
        hipLaunchKernelGGL((k_tri_seq_re<R, C, DIM == 2, HelmProb<double, Ti, To, DIM, REAL>>), dim3((unsigned)blocks),
                           dim3(NT), 0, st, pb, t.foff, t.fmode, (const double *)t.fac, t.nent, t.nfx, t.batch,
                           pb.bands.nop);
    } else {
        const long blocks = (items + C - 1) / C;
        if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_tri_seq<R, C, DIM == 2, HelmProb<double, Ti, To, DIM, REAL>, OT>), dim3((unsigned)blocks),
                           dim3(NT), 0, st, pb, t.foff, t.fmode, (const OT *)t.fac, t.nent, t.nfx, t.batch, pb.bands.nop);
    }
    return hipGetLastError();
}

// As many chains per block as 512 threads hold with R rows per lane.
template <int R, typename Ti, typename To, int DIM, bool REAL>
static hipError_t launch_seq_R(const TriHelm &t, const void *in, void *out, hipStream_t st) {
    const int n = t.dim == 2 ? t.Ny : t.Nz;
    const int g = (n + R - 1) / R;   // positions of one chain
    auto fits = [&](int C) { const int PW = 64 / C; return ((g + PW - 1) / PW * PW) * C <= 512; };
    if (fits(64)) return launch_seq_rc<R, 64, Ti, To, DIM, REAL>(t, in, out, st);
    if (fits(32)) return launch_seq_rc<R, 32, Ti, To, DIM, REAL>(t, in, out, st);
    if (fits(16)) return launch_seq_rc<R, 16, Ti, To, DIM, REAL>(t, in, out, st);
    if (fits(8)) return launch_seq_rc<R, 8, Ti, To, DIM, REAL>(t, in, out, st);
    if (fits(4)) return launch_seq_rc<R, 4, Ti, To, DIM, REAL>(t, in, out, st);
    if (fits(2)) return launch_seq_rc<R, 2, Ti, To, DIM, REAL>(t, in, out, st);
    if (fits(1)) return launch_seq_rc<R, 1, Ti, To, DIM, REAL>(t, in, out, st);
    return hipErrorInvalidValue;
}

// 16 rows per lane with real factors; 8 with complex ones (twice the
// registers), except above n = 4096 where 512 threads need 16
template <typename Ti, typename To, int DIM, bool REAL>
static hipError_t launch_seq_d(const TriHelm &t, const void *in, void *out, hipStream_t st) {
    const int n = t.dim == 2 ? t.Ny : t.Nz;
    if (REAL || n > 4096) return launch_seq_R<16, Ti, To, DIM, REAL>(t, in, out, st);
    return launch_seq_R<8, Ti, To, DIM, REAL>(t, in, out, st);
}

template <typename Ti, typename To>
static hipError_t launch_seq_t(const TriHelm &t, const void *in, void *out, hipStream_t st) {
    if (t.dim == 2) return t.real ? launch_seq_d<Ti, To, 2, true>(t, in, out, st) : launch_seq_d<Ti, To, 2, false>(t, in, out, st);
    return t.real ? launch_seq_d<Ti, To, 3, true>(t, in, out, st) : launch_seq_d<Ti, To, 3, false>(t, in, out, st);
}

hipError_t launch_tri_seq(const TriHelm &t, DstTypes ty, const void *in, void *out, hipStream_t st) {
    if (t.nfx <= 0) return hipSuccess;
    if (!t.foff || !t.fmode || !t.fac || t.batch < 1) return hipErrorInvalidValue;
    if (ty.in == 0 && ty.out == 0) return launch_seq_t<double, double>(t, in, out, st);
    if (ty.in == 1 && ty.out == 1) return launch_seq_t<float, float>(t, in, out, st);
    return hipErrorInvalidValue;
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
