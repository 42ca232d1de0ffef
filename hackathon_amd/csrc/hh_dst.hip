// Batched orthonormal DST-I kernels for gfx950.
//
// Replaces the reference's FFTW / cuFFT transforms:
//   2D_Program/solver_functions/DST.c:6-44 (forwardDST / reverseDST, one
//     r2c FFT of length 2Nx+2 per row and per real/imaginary part)
//   3D_Program/solver_functions/DST.c:13-204 and cuda_kernels.cu:23-176
//     (2D DST of each z layer: rows, transpose, columns).
// X_k = sqrt(2/(N+1)) sum_n x_n sin(pi n k/(N+1)) is taken from the FFT of
// the odd extension z = [0, x, 0, -reverse(x)] (length L = 2(N+1)):
// X_k = sqrt(2/(N+1)) (i/2) Z_k.  Real and imaginary parts are transformed
// together as one complex line.
//
// Kernels:
//   k_dst_fft_rows  lines contiguous in memory (x axis); inputs read straight
//                   from HBM into registers for the first radix-16 stage,
//                   outputs stored straight from the last stage (both
//                   coalesced), LDS only between stages.
//   k_dst_fft_cols  lines strided (y axis of a 3D layer): a tile of LPB
//                   neighbouring columns is staged through LDS so that HBM
//                   is read and written in LPB*16-byte row segments.
//   k_dst_gemm      any N (e.g. N = 100): dense DST matrix times the lines.
#include "hh_common.h"
#include "hh_fft.h"
#include "hh_internal.h"

namespace hh {

template <typename T> struct tsel;
template <> struct tsel<double> { static constexpr int id = 0; };
template <> struct tsel<float> { static constexpr int id = 1; };

// TPB: target threads per block (256, or 64 for one-wave blocks whose
// barriers are nearly free and whose small LDS footprint lets more blocks
// share a CU)
// LDS line stride (doubles) of a block's line buffers [re | im]: at least
// 2 LP, and congruent to 32/LPB modulo 32 so that the tile loads and stores
// of the strided kernels (lanes across the block's lines, then along them)
// fall on distinct banks.
constexpr int line_stride(int LP, int LPB) {
    const int want = LPB >= 32 ? 1 : 32 / LPB;
    int ls = 2 * LP;
    while (ls % 32 != want % 32) ls++;
    return ls;
}

template <int L, int TPB = 256> struct dst_cfg {
    static constexpr int P = fft_shape<L>::P;
    static constexpr int LPB = P >= TPB ? 1 : TPB / P;   // lines per block
    static constexpr int THREADS = P * LPB;
    static constexpr int LP = fft_shape<L>::LP;
    static constexpr int LS = line_stride(LP, LPB);
};

// Position of element k of line `row` (the lines of every right-hand side
// stacked, Ny lines each) in a buffer that is row-major (tw = 0) or
// mode-tiled (tw > 0, a power of two): the N modes of a right-hand side are
// cut into tiles of tw neighbouring modes, tile t0 holds its modes for all
// Ny lines contiguously, [t0 * Ny + j * width + (k - t0)], width = min(tw,
// N - t0).  The tiled form is what the tridiagonal pass reads: each of its
// threads walks down a tile in whole width*16-byte segments.
struct LinePos {
    long base;
    int j, N, Ny, tw;
    HH_DEV long at(int k) const {
        if (tw == 0) return base + k;
        const int t0 = k & ~(tw - 1);
        const int w = min(tw, N - t0);
        return base + (long)t0 * Ny + (long)j * w + (k - t0);
    }
};
struct RowMap {
    int N, Ny, tw;
    long field;
    double acc;   // output map: nonzero = accumulate (see Lines::acc)
    HH_DEV LinePos line(long row) const {
        LinePos p;
        p.N = N; p.Ny = Ny; p.tw = tw;
        if (tw == 0) { p.base = row * N; p.j = 0; return p; }
        const long b = row / Ny;
        p.j = (int)(row - b * Ny);
        p.base = b * field;
        return p;
    }
    // the line `l` rows after a known one (l small: steps over right-hand sides by subtraction)
    HH_DEV LinePos next(LinePos p, int l) const {
        if (tw == 0) { p.base += (long)l * N; return p; }
        int j = p.j + l;
        while (j >= Ny) { j -= Ny; p.base += field; }
        p.j = j;
        return p;
    }
};

// Store of one output element: out = X, or out += acc X for an FP64 output
// with acc != 0 (the refinement update folded into the last pass).
template <typename To, typename T> HH_DEV void put_out(C2<To> *o, C2<T> v, double acc) {
    if constexpr (sizeof(To) == 8) {
        if (acc != 0) {
            const C2<double> u = *o;
            *o = mk<double>(u.x + acc * (double)v.x, u.y + acc * (double)v.y);
            return;
        }
    }
    *o = cconv<To, T>(v);
}

// Coalesced load of LPB consecutive lines from the mode-tiled layout into
// the LDS line buffers (element k of line l at position k+1), in tile-major
// order so that each tile's LPB x tw elements are read as one contiguous
// run.  Every load of a thread is issued before the first LDS write.
template <typename Ti, typename T, int N, int LPB, int LP, int LS, int NT>
HH_DEV void load_tiled_rows(T *smem, const C2<Ti> *in, const RowMap &map, long row0, int nr) {
    constexpr int LGL = LPB == 1 ? 0 : LPB == 2 ? 1 : LPB == 4 ? 2 : LPB == 8 ? 3 : LPB == 16 ? 4
                      : LPB == 32 ? 5 : LPB == 64 ? 6 : LPB == 128 ? 7 : 8;
    static_assert((1 << LGL) == LPB, "lines per block must be a power of two");
    constexpr int ITER = ((N + 63) * LPB + NT - 1) / NT;   // tiles are at most 64 wide
    const int lg = 31 - __clz(map.tw);
    const int tw = 1 << lg;
    const int slots = ((N + tw - 1) >> lg) << (lg + LGL);
    const LinePos l0 = map.line(row0);
    C2<T> v[ITER];
    int pos[ITER];
#pragma unroll
    for (int i = 0; i < ITER; i++) {
        const int e = threadIdx.x + i * NT;
        const int tile = e >> (lg + LGL);
        const int rem = e & ((1 << (lg + LGL)) - 1);
        const int l = rem >> lg, k = (tile << lg) + (rem & (tw - 1));
        const bool ok = e < slots && k < N && l < nr;
        v[i] = cconv<T, Ti>(in[ok ? map.next(l0, l).at(k) : l0.at(0)]);
        pos[i] = ok ? l * LS + lds_pad(k + 1) : -1;
    }
#pragma unroll
    for (int i = 0; i < ITER; i++) {
        if (pos[i] >= 0) {
            smem[pos[i]] = v[i].x;
            smem[pos[i] + LP] = v[i].y;
        }
    }
}

template <typename Ti, typename T, typename To, int L, int TPB, bool TIN>
__global__ __launch_bounds__((dst_cfg<L, TPB>::THREADS)) void k_dst_fft_rows(
    const C2<Ti> *in, C2<To> *out, long nrows, T scale,
    const C2<T> *__restrict__ tw, RowMap min_, RowMap mout) {
    constexpr int P = dst_cfg<L, TPB>::P, LPB = dst_cfg<L, TPB>::LPB, LP = dst_cfg<L, TPB>::LP,
                  LS = dst_cfg<L, TPB>::LS;
    constexpr int M = L / 2, N = M - 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    const int li = threadIdx.x / P, t = threadIdx.x % P;
    T *sre = smem + li * LS, *sim = smem + li * LS + LP;
    const long row = (long)blockIdx.x * LPB + li;
    const bool valid = row < nrows;
    const LinePos lout = mout.line(valid ? row : 0);
    if constexpr (TIN) {   // tiled input: stage the block's lines through LDS
        constexpr int NT = dst_cfg<L, TPB>::THREADS;
        const long row0 = (long)blockIdx.x * LPB;
        const long nr = nrows - row0 < LPB ? nrows - row0 : LPB;
        load_tiled_rows<Ti, T, N, LPB, LP, LS, NT>(smem, in, min_, row0, (int)nr);
        __syncthreads();
    }
    // Branch-free: every load is issued (index clamped into the row) and
    // the zero entries of the odd extension are selected afterwards, so the
    // 16 loads of a thread are all in flight together.
    auto get = [&](int n, int) -> C2<T> {
        const bool zero = (n == 0) || (n == M);
        const int i = zero ? 0 : (n < M ? n - 1 : L - n - 1);
        C2<T> v;
        if constexpr (TIN) v = mk<T>(sre[lds_pad(i + 1)], sim[lds_pad(i + 1)]);
        else v = cconv<T, Ti>(in[(valid ? row : 0) * N + i]);   // row-major input
        const C2<T> r = n < M ? v : cneg<T>(v);
        return zero ? mk<T>(0, 0) : r;
    };
    auto put = [&](int k, C2<T> z) {
        if (valid && k >= 1 && k <= N) put_out<To, T>(out + lout.at(k - 1), mk<T>(-z.y * scale, z.x * scale), mout.acc);
    };
    fft_line<T, L, false>(sre, sim, tw, t, get, put);
}

// Coalesced load of a tile of LPB neighbouring strided lines into the LDS
// line buffers (element q of line c at position q+1).  Every load of a
// thread is issued before the first LDS write so that they overlap; lanes
// past the end of the tile read the tile's first element and write nothing.
template <typename Ti, typename T, int N, int LPB, int LP, int LS, int NT>
HH_DEV void load_col_tile(T *smem, const C2<Ti> *in, long base, long c0, long ncols, long elem_stride) {
    constexpr int ITER = (N * LPB + NT - 1) / NT;
    C2<T> v[ITER];
#pragma unroll
    for (int i = 0; i < ITER; i++) {
        const int e = threadIdx.x + i * NT;
        const int c = e % LPB, q = e / LPB;
        const bool ok = (e < N * LPB) && (c0 + c < ncols);
        v[i] = cconv<T, Ti>(in[ok ? base + c + (long)q * elem_stride : base]);
        if (!ok) v[i] = mk<T>(0, 0);
    }
#pragma unroll
    for (int i = 0; i < ITER; i++) {
        const int e = threadIdx.x + i * NT;
        if (e < N * LPB) {
            const int c = e % LPB, q = e / LPB;
            smem[c * LS + lds_pad(q + 1)] = v[i].x;
            smem[c * LS + LP + lds_pad(q + 1)] = v[i].y;
        }
    }
}

// Store of the tile back to its strided lines (positions 1..N of each line).
template <typename T, typename To, int N, int LPB, int LP, int LS, int NT>
HH_DEV void store_col_tile(const T *smem, C2<To> *out, long base, long c0, long ncols, long elem_stride) {
    constexpr int ITER = (N * LPB + NT - 1) / NT;
#pragma unroll
    for (int i = 0; i < ITER; i++) {
        const int e = threadIdx.x + i * NT;
        const int c = e % LPB, q = e / LPB;
        if (e < N * LPB && c0 + c < ncols)
            out[base + c + (long)q * elem_stride] =
                cconv<To, T>(mk<T>(smem[c * LS + lds_pad(q + 1)], smem[c * LS + LP + lds_pad(q + 1)]));
    }
}

template <typename Ti, typename T, typename To, int L, int TPB>
__global__ __launch_bounds__((dst_cfg<L, TPB>::THREADS)) void k_dst_fft_cols(
    const C2<Ti> *in, C2<To> *out, long ncols, long ntiles,
    long plane_stride, long elem_stride, T scale, const C2<T> *__restrict__ tw) {
    constexpr int P = dst_cfg<L, TPB>::P, LPB = dst_cfg<L, TPB>::LPB, LP = dst_cfg<L, TPB>::LP,
                  LS = dst_cfg<L, TPB>::LS;
    constexpr int M = L / 2, N = M - 1, NT = dst_cfg<L, TPB>::THREADS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    const long tile = blockIdx.x % ntiles, plane = blockIdx.x / ntiles;
    const long c0 = tile * LPB;
    const long base = plane * plane_stride + c0;
    // coalesced tile load: lanes run along the columns of one row
    load_col_tile<Ti, T, N, LPB, LP, LS, NT>(smem, in, base, c0, ncols, elem_stride);
    __syncthreads();
    const int li = threadIdx.x / P, t = threadIdx.x % P;
    T *sre = smem + li * LS, *sim = smem + li * LS + LP;
    auto get = [&](int n, int) -> C2<T> {
        if (n == 0 || n == M) return mk<T>(0, 0);
        if (n < M) return mk<T>(sre[lds_pad(n)], sim[lds_pad(n)]);
        return mk<T>(-sre[lds_pad(L - n)], -sim[lds_pad(L - n)]);
    };
    auto put = [&](int k, C2<T> z) {
        if (k >= 1 && k <= N) {
            sre[lds_pad(k)] = -z.y * scale;
            sim[lds_pad(k)] = z.x * scale;
        }
    };
    fft_line<T, L, true>(sre, sim, tw, t, get, put);
    __syncthreads();
    store_col_tile<T, To, N, LPB, LP, LS, NT>(smem, out, base, c0, ncols, elem_stride);
}

// ------------------------------------------------------------------------
// Half-length DST-I.  With M = N+1 and x_0 = x_M = 0, the row
//   w_j = sin(pi j/M)(x_j + x_{M-j}) + (x_j - x_{M-j})/2,  j = 1..M-1, w_0 = 0
// has the M-point transform F_k = sum_j w_j e^{-2 pi i jk/M} with
//   S_{2k}   = (i/2)(F_k - F_{M-k})
//   S_{2k+1} = S_{2k-1} + (F_k + F_{M-k})/2,   S_{-1} = -S_1
// where S_m = sum_j x_j sin(pi jm/M).  (The symmetric part of w carries
// cos-weighted differences of odd outputs, the antisymmetric part the even
// outputs; real and imaginary parts of x ride through the same complex
// transform because every weight is real.)  This halves the FFT length and
// the LDS of the odd-extension transform; the odd outputs need a prefix sum
// over M/2 terms, done with wave shuffles inside a line and through LDS
// across the waves of a line.
template <int M, int TPB = 256> struct half_cfg {
    static constexpr int P = fft_shape<M>::P;
    static constexpr int LPB = P >= TPB ? 1 : TPB / P;
    static constexpr int THREADS = P * LPB;
    static constexpr int LP = fft_shape<M>::LP;
    static constexpr int LS = line_stride(LP, LPB);
    static constexpr int K = (M / 2) / P;   // pairs (F_k, F_{M-k}) per thread, M >= 2P
};

// sin(pi n/M) of the first-stage value n = t + r P of lane t:
// sin(pi t/M) cos(pi r P/M) + cos(pi t/M) sin(pi r P/M), the lane's pair
// loaded once and the slot's pair the same in every lane (uniform loads),
// from sn[j] = sin(pi j/M), j < M, with cos(pi a/M) = +-sn[|M/2 - a|].
template <typename T, int M> struct HalfSine {
    const T *sn;
    T st, ct;
    HH_DEV HalfSine(const T *s, int t) : sn(s), st(s[t]), ct(s[M / 2 - t]) {}
    HH_DEV T at(int r) const {
        const int a = r * fft_shape<M>::P;
        const T sr = sn[a];
        const T cr = a <= M / 2 ? sn[M / 2 - a] : -sn[a - M / 2];
        return st * cr + ct * sr;
    }
};

// Post-processing of one line whose transform F sits in (sre, sim); writes the
// scaled outputs X_m = sqrt(2/M) S_m (m = 1..M-1) back to positions m of the
// same buffer.  tot: LDS scratch of one complex per wave of the line.
template <typename T, int M>
HH_DEV void half_post(T *sre, T *sim, T *tot_re, T *tot_im, int t, T scale) {
    constexpr int P = half_cfg<M>::P, K = half_cfg<M>::K;
    C2<T> se[K], so[K];
    C2<T> run = mk<T>(0, 0);
#pragma unroll
    for (int u = 0; u < K; u++) {
        const int k = t * K + u;
        const int km = (M - k) & (M - 1);
        const C2<T> fk = mk<T>(sre[lds_pad(k)], sim[lds_pad(k)]);
        const C2<T> fm = mk<T>(sre[lds_pad(km)], sim[lds_pad(km)]);
        const C2<T> dif = csub<T>(fk, fm);
        se[u] = mk<T>(-(T)0.5 * dif.y, (T)0.5 * dif.x);            // (i/2)(F_k - F_{M-k})
        run = cadd<T>(run, cscale<T>(cadd<T>(fk, fm), (T)0.5));   // running sum of D_k
        so[u] = run;                                              // inclusive prefix within the thread
    }
    // exclusive scan of the thread totals across the line
    C2<T> excl = mk<T>(0, 0);
    {
        constexpr int SEG = P < 64 ? P : 64;
        const int lane = t % SEG;
        C2<T> inc = run;
#pragma unroll
        for (int d = 1; d < SEG; d <<= 1) {
            T ox = __shfl_up(inc.x, d, SEG), oy = __shfl_up(inc.y, d, SEG);
            if (lane >= d) { inc.x += ox; inc.y += oy; }
        }
        excl = csub<T>(inc, run);
        if constexpr (P > 64) {
            const int wv = t / 64;
            if (lane == 63) { tot_re[wv] = inc.x; tot_im[wv] = inc.y; }
            __syncthreads();
            for (int q = 0; q < wv; q++) { excl.x += tot_re[q]; excl.y += tot_im[q]; }
        }
    }
    // D_0 / 2 = S_1: the first thread's first D is F_0 (F_M = F_0)
    const C2<T> f0 = mk<T>(sre[lds_pad(0)], sim[lds_pad(0)]);
    const C2<T> s1 = cscale<T>(f0, (T)0.5);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < K; u++) {
        const int k = t * K + u;
        // S_{2k+1} = S_{-1} + sum_{k' <= k} D_k' = -S_1 + prefix
        const C2<T> sodd = csub<T>(cadd<T>(so[u], excl), s1);
        if (k > 0) {
            sre[lds_pad(2 * k)] = se[u].x * scale;
            sim[lds_pad(2 * k)] = se[u].y * scale;
        }
        sre[lds_pad(2 * k + 1)] = sodd.x * scale;
        sim[lds_pad(2 * k + 1)] = sodd.y * scale;
    }
    __syncthreads();
}

// rows (contiguous lines); aux = [twiddles of length M | sin(pi j/M), j < M]
template <typename Ti, typename T, typename To, int M, int TPB, bool TIN>
__global__ __launch_bounds__((half_cfg<M, TPB>::THREADS)) void k_dst_half_rows(const C2<Ti> *in, C2<To> *out,
                                                                            long nrows, T scale,
                                                                            const C2<T> *__restrict__ aux,
                                                                            RowMap min_, RowMap mout) {
    using cfg = half_cfg<M, TPB>;
    constexpr int P = cfg::P, LPB = cfg::LPB, LP = cfg::LP, LS = cfg::LS, N = M - 1, NT = cfg::THREADS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    T *tot = smem + LPB * LS;   // per line: 2 * (P/64 + 1) reals
    const int li = threadIdx.x / P, t = threadIdx.x % P;
    T *sre = smem + li * LS, *sim = smem + li * LS + LP;
    const long row0 = (long)blockIdx.x * LPB;
    const long row = row0 + li;
    const bool valid = row < nrows;
    const T *sn = reinterpret_cast<const T *>(aux + M);
    const long nr = nrows - row0 < LPB ? nrows - row0 : LPB;
    if constexpr (TIN) {   // tiled input: stage the block's lines through LDS
        load_tiled_rows<Ti, T, N, LPB, LP, LS, NT>(smem, in, min_, row0, (int)nr);
        __syncthreads();
    }
    // branch-free (clamped indices, zero selected for n = 0) so that all
    // loads of a thread are in flight together; rows past the end read row 0
    const HalfSine<T, M> hs(sn, t);
    auto get = [&](int n, int r) -> C2<T> {
        const int ia = n > 0 ? n - 1 : 0, ib = n > 0 ? M - n - 1 : 0;
        C2<T> a, b;
        if constexpr (TIN) {
            a = mk<T>(sre[lds_pad(ia + 1)], sim[lds_pad(ia + 1)]);
            b = mk<T>(sre[lds_pad(ib + 1)], sim[lds_pad(ib + 1)]);
        } else {   // row-major input
            const C2<Ti> *x = in + (valid ? row : 0) * N;
            a = cconv<T, Ti>(x[ia]);
            b = cconv<T, Ti>(x[ib]);
        }
        const T s = hs.at(r);
        const C2<T> z = mk<T>(s * (a.x + b.x) + (T)0.5 * (a.x - b.x), s * (a.y + b.y) + (T)0.5 * (a.y - b.y));
        return n > 0 ? z : mk<T>(0, 0);
    };
    auto put = [&](int k, C2<T> z) {
        sre[lds_pad(k)] = z.x;
        sim[lds_pad(k)] = z.y;
    };
    fft_line<T, M, true>(sre, sim, aux, t, get, put);
    __syncthreads();
    constexpr int NW = P > 64 ? P / 64 : 1;
    half_post<T, M>(sre, sim, tot + li * 2 * NW, tot + li * 2 * NW + NW, t, scale);
    // coalesced store of the block's rows (contiguous in memory)
    constexpr int ITER = (LPB * N + NT - 1) / NT;
    if (mout.tw == 0) {
        C2<To> *y = out + row0 * N;
#pragma unroll
        for (int i = 0; i < ITER; i++) {
            const int e = threadIdx.x + i * NT;
            if (e < nr * N) {
                const int l = e / N, m = e % N + 1;
                put_out<To, T>(y + e, mk<T>(smem[l * LS + lds_pad(m)], smem[l * LS + LP + lds_pad(m)]), mout.acc);
            }
        }
    } else {
        // tile-major order: the block's lines of one tile are contiguous there
        const int lg = 31 - __clz(mout.tw);   // tw is a power of two, and so is LPB
        constexpr int LGL = LPB == 1 ? 0 : LPB == 2 ? 1 : LPB == 4 ? 2 : LPB == 8 ? 3 : LPB == 16 ? 4
                          : LPB == 32 ? 5 : LPB == 64 ? 6 : LPB == 128 ? 7 : 8;
        static_assert((1 << LGL) == LPB, "lines per block must be a power of two");
        const int tw = 1 << lg;
        const int slots = ((N + tw - 1) >> lg) << (lg + LGL);
        // the block's lines, each with its position in the tiled field
        const LinePos l0 = mout.line(row0);
        for (int e = threadIdx.x; e < slots; e += NT) {
            const int tile = e >> (lg + LGL);
            const int rem = e & ((1 << (lg + LGL)) - 1);
            const int l = rem >> lg, k = (tile << lg) + (rem & (tw - 1));
            if (k < N && l < nr)
                put_out<To, T>(out + mout.next(l0, l).at(k),
                               mk<T>(smem[l * LS + lds_pad(k + 1)], smem[l * LS + LP + lds_pad(k + 1)]), mout.acc);
        }
    }
}

template <typename Ti, typename T, typename To, int M, int TPB>
__global__ __launch_bounds__((half_cfg<M, TPB>::THREADS)) void k_dst_half_cols(const C2<Ti> *in, C2<To> *out,
                                                                            long ncols, long ntiles, long plane_stride,
                                                                            long elem_stride, T scale,
                                                                            const C2<T> *__restrict__ aux) {
    using cfg = half_cfg<M, TPB>;
    constexpr int P = cfg::P, LPB = cfg::LPB, LP = cfg::LP, LS = cfg::LS, N = M - 1, NT = cfg::THREADS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    T *tot = smem + LPB * LS;
    const long tile = blockIdx.x % ntiles, plane = blockIdx.x / ntiles;
    const long c0 = tile * LPB;
    const long base = plane * plane_stride + c0;
    load_col_tile<Ti, T, N, LPB, LP, LS, NT>(smem, in, base, c0, ncols, elem_stride);
    __syncthreads();
    const int li = threadIdx.x / P, t = threadIdx.x % P;
    T *sre = smem + li * LS, *sim = smem + li * LS + LP;
    const T *sn = reinterpret_cast<const T *>(aux + M);
    const HalfSine<T, M> hs(sn, t);
    auto get = [&](int n, int r) -> C2<T> {
        if (n == 0) return mk<T>(0, 0);
        const int nb = M - n;
        const T ar = sre[lds_pad(n)], ai = sim[lds_pad(n)];
        const T br = nb < M ? sre[lds_pad(nb)] : (T)0, bi = nb < M ? sim[lds_pad(nb)] : (T)0;
        const T s = hs.at(r);
        return mk<T>(s * (ar + br) + (T)0.5 * (ar - br), s * (ai + bi) + (T)0.5 * (ai - bi));
    };
    auto put = [&](int k, C2<T> z) {
        sre[lds_pad(k)] = z.x;
        sim[lds_pad(k)] = z.y;
    };
    fft_line<T, M, true>(sre, sim, aux, t, get, put);
    __syncthreads();
    constexpr int NW = P > 64 ? P / 64 : 1;
    half_post<T, M>(sre, sim, tot + li * 2 * NW, tot + li * 2 * NW + NW, t, scale);
    store_col_tile<T, To, N, LPB, LP, LS, NT>(smem, out, base, c0, ncols, elem_stride);
}

// Dense DST: out[line][k] = sum_n S[k][n] in[line][n].  Block tile: 32 lines x
// 64 outputs, 256 threads, each thread 8 outputs of one line.
template <typename Ti, typename T, typename To>
__global__ __launch_bounds__(256) void k_dst_gemm(const C2<Ti> *__restrict__ in, C2<To> *__restrict__ out,
                                                  Lines ln, long nlines, const T *__restrict__ S) {
    constexpr int TL = 32, TK = 64, TN = 32;
    __shared__ C2<T> xs[TL][TN + 1];
    __shared__ T ss[TN][TK + 1];
    const int N = ln.N;
    const long l0 = (long)blockIdx.x * TL;
    const int k0 = blockIdx.y * TK;
    const int tl = threadIdx.x % TL, tk = threadIdx.x / TL;  // tk in [0,8)
    C2<T> acc[8];
#pragma unroll
    for (int u = 0; u < 8; u++) acc[u] = mk<T>(0, 0);
    auto line_base = [&](long l) -> long {
        long p = l / ln.ncols, c = l % ln.ncols;
        return p * ln.plane_stride + c * ln.col_stride;
    };
    for (int n0 = 0; n0 < N; n0 += TN) {
        for (int e = threadIdx.x; e < TL * TN; e += 256) {
            int a, b;  // a: line in tile, b: n in tile
            if (ln.elem_stride == 1) { a = e / TN; b = e % TN; }
            else { a = e % TL; b = e / TL; }
            long l = l0 + a;
            int n = n0 + b;
            C2<T> v = mk<T>(0, 0);
            if (l < nlines && n < N) v = cconv<T, Ti>(in[line_base(l) + (long)n * ln.elem_stride]);
            xs[a][b] = v;
        }
        for (int e = threadIdx.x; e < TN * TK; e += 256) {
            int b = e / TK, kk = e % TK;
            int n = n0 + b, k = k0 + kk;
            ss[b][kk] = (n < N && k < N) ? S[(long)n * N + k] : (T)0;
        }
        __syncthreads();
#pragma unroll 4
        for (int b = 0; b < TN; b++) {
            C2<T> xv = xs[tl][b];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                T s = ss[b][tk + 8 * u];
                acc[u].x += xv.x * s;
                acc[u].y += xv.y * s;
            }
        }
        __syncthreads();
    }
    long l = l0 + tl;
    if (l < nlines) {
        long bse = line_base(l);
#pragma unroll
        for (int u = 0; u < 8; u++) {
            int k = k0 + tk + 8 * u;
            if (k < N) put_out<To, T>(out + bse + (long)k * ln.elem_stride, acc[u], ln.acc);
        }
    }
}

// Dense DST on the matrix cores: the lines' real and imaginary parts form
// the rows of a real matrix X (row 2l + c = part c of line l), and
// OUT = X S is a real GEMM.  Block tile 64 rows (32 lines) x 64 outputs,
// four waves of 2 x 2 MFMA 16x16 tiles (v_mfma_f64_16x16x4f64 in FP64,
// v_mfma_f32_16x16x4f32 in FP32, FP32 accumulation), K staged through LDS 16
// at a time, the next 16 loaded into registers meanwhile; the output tile goes back through LDS so that stores are
// coalesced for contiguous and for strided lines.  Operand layouts (lane l):
// A[l % 16][l / 16], B[l / 16][l % 16]; result j of lane l is row
// l / 16 + 4 j (FP64) or 4 (l / 16) + j (FP32) of column l % 16.
template <typename T> struct Mfma16;
template <> struct Mfma16<double> {
    typedef double V __attribute__((ext_vector_type(4)));
    static HH_DEV V step(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static HH_DEV int row(int l, int j) { return l / 16 + 4 * j; }
};
template <> struct Mfma16<float> {
    typedef float V __attribute__((ext_vector_type(4)));
    static HH_DEV V step(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static HH_DEV int row(int l, int j) { return 4 * (l / 16) + j; }
};

template <typename Ti, typename T, typename To>
__global__ __launch_bounds__(256) void k_dst_mfma(const C2<Ti> *__restrict__ in, C2<To> *__restrict__ out, Lines ln,
                                                  long nlines, const T *__restrict__ S) {
    using M = Mfma16<T>;
    constexpr int BL = 32, BM = 2 * BL, BN = 64, BK = 16, PAD = 1;
    // staging [k][m] and [k][n] during the product, the output tile [m][n] after it
    constexpr int STAGE = BK * (BM + PAD) + BK * (BN + PAD), OUTT = BM * (BN + PAD);
    __shared__ T sh[STAGE > OUTT ? STAGE : OUTT];
    T *xs = sh, *ss = sh + BK * (BM + PAD);
    const int N = ln.N;
    const long l0 = (long)blockIdx.x * BL;
    const int n0 = blockIdx.y * BN;
    const int t = threadIdx.x, lane = t % 64, wv = t / 64;
    const int wm = (wv % 2) * 32, wn = (wv / 2) * 32;   // the wave's 32 x 32 sub-tile
    auto line_base = [&](long l) -> long {
        const long p = l / ln.ncols, c = l % ln.ncols;
        return p * ln.plane_stride + c * ln.col_stride;
    };
    typename M::V acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = typename M::V{0, 0, 0, 0};
    // chunk k0 of X (BL lines x BK elements; contiguous lines: threads along
    // the line, strided lines: threads across lines) and of S, into
    // registers: the next chunk's loads are in flight while the matrix cores
    // work on the current one
    constexpr int XI = BL * BK / 256, SI = BK * BN / 256;
    C2<T> xv[XI];
    T sv[SI];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < XI; i++) {
            const int e = t + 256 * i;
            int a, b;
            if (ln.elem_stride == 1) { a = e / BK; b = e % BK; }
            else { a = e % BL; b = e / BL; }
            const long l = l0 + a;
            const int k = k0 + b;
            xv[i] = (l < nlines && k < N) ? cconv<T, Ti>(in[line_base(l) + (long)k * ln.elem_stride]) : mk<T>(0, 0);
        }
#pragma unroll
        for (int i = 0; i < SI; i++) {
            const int e = t + 256 * i;
            const int b = e / BN, c = e % BN;
            const int k = k0 + b, n = n0 + c;
            sv[i] = (k < N && n < N) ? S[(long)k * N + n] : (T)0;
        }
    };
    load(0);
    for (int k0 = 0; k0 < N; k0 += BK) {
#pragma unroll
        for (int i = 0; i < XI; i++) {
            const int e = t + 256 * i;
            int a, b;
            if (ln.elem_stride == 1) { a = e / BK; b = e % BK; }
            else { a = e % BL; b = e / BL; }
            xs[b * (BM + PAD) + 2 * a] = xv[i].x;
            xs[b * (BM + PAD) + 2 * a + 1] = xv[i].y;
        }
#pragma unroll
        for (int i = 0; i < SI; i++) {
            const int e = t + 256 * i;
            ss[(e / BN) * (BN + PAD) + e % BN] = sv[i];
        }
        __syncthreads();
        if (k0 + BK < N) load(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + lane / 16;
            T a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++) a[i] = xs[kr * (BM + PAD) + wm + 16 * i + lane % 16];
#pragma unroll
            for (int j = 0; j < 2; j++) b[j] = ss[kr * (BN + PAD) + wn + 16 * j + lane % 16];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = M::step(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    // output tile through LDS
    T *os = sh;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                os[(wm + 16 * i + M::row(lane, q)) * (BN + PAD) + wn + 16 * j + lane % 16] = acc[i][j][q];
    __syncthreads();
#pragma unroll
    for (int e = t; e < BL * BN; e += 256) {
        int a, c;
        if (ln.elem_stride == 1) { a = e / BN; c = e % BN; }
        else { a = e % BL; c = e / BL; }
        const long l = l0 + a;
        const int n = n0 + c;
        if (l < nlines && n < N)
            put_out<To, T>(out + line_base(l) + (long)n * ln.elem_stride,
                           mk<T>(os[2 * a * (BN + PAD) + c], os[(2 * a + 1) * (BN + PAD) + c]), ln.acc);
    }
}

bool fft_supported(int N) {
    int L = 2 * (N + 1);
    return N >= 1 && L <= HH_FFT_LMAX && (L & (L - 1)) == 0 && L >= 4;
}

template <typename Ti, typename T, typename To, int L, int TPB>
static hipError_t launch_fft_L(const Lines &ln, const void *in, void *out, const void *tw, hipStream_t st) {
    using cfg = dst_cfg<L, TPB>;
    size_t shm = sizeof(T) * cfg::LPB * cfg::LS;
    T scale = (T)sqrt(0.5 / (double)(L / 2));
    const C2<Ti> *pi = (const C2<Ti> *)in;
    C2<To> *po = (C2<To> *)out;
    const C2<T> *ptw = (const C2<T> *)tw;
    if (ln.elem_stride == 1 && ln.col_stride == ln.N && ln.nplanes * ln.ncols > 0 &&
        (ln.nplanes == 1 || ln.plane_stride == ln.ncols * ln.N)) {
        long nrows = ln.ncols * ln.nplanes;
        auto kfn = ln.in_tw ? k_dst_fft_rows<Ti, T, To, L, TPB, true> : k_dst_fft_rows<Ti, T, To, L, TPB, false>;
        if (shm > 65536) {
            hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            if (e != hipSuccess) return e;
        }
        long grid = (nrows + cfg::LPB - 1) / cfg::LPB;
        const RowMap rin{ln.N, ln.ny, ln.in_tw, ln.field, 0.0}, rout{ln.N, ln.ny, ln.out_tw, ln.field, ln.acc};
        hipLaunchKernelGGL(kfn, dim3((unsigned)grid), dim3(cfg::THREADS), shm, st, pi, po, nrows, scale, ptw, rin,
                           rout);
        return hipGetLastError();
    }
    if (ln.col_stride != 1 || ln.acc != 0) return hipErrorInvalidValue;   // strided lines: plain stores
    auto kfn = k_dst_fft_cols<Ti, T, To, L, TPB>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    long ntiles = (ln.ncols + cfg::LPB - 1) / cfg::LPB;
    long grid = ntiles * ln.nplanes;
    hipLaunchKernelGGL(kfn, dim3((unsigned)grid), dim3(cfg::THREADS), shm, st, pi, po, ln.ncols, ntiles,
                       ln.plane_stride, ln.elem_stride, scale, ptw);
    return hipGetLastError();
}

template <typename Ti, typename T, typename To, int M, int TPB>
static hipError_t launch_half_M(const Lines &ln, const void *in, void *out, const void *aux, bool wide,
                                hipStream_t st) {
    using cfg = half_cfg<M, TPB>;
    constexpr int NW = cfg::P > 64 ? cfg::P / 64 : 1;
    size_t shm = sizeof(T) * (cfg::LPB * cfg::LS + 2 * cfg::LPB * NW);
    T scale = (T)sqrt(2.0 / (double)M);
    const C2<Ti> *pi = (const C2<Ti> *)in;
    C2<To> *po = (C2<To> *)out;
    const C2<T> *pa = (const C2<T> *)aux;
    if (ln.elem_stride == 1 && ln.col_stride == ln.N && (ln.nplanes == 1 || ln.plane_stride == ln.ncols * ln.N)) {
        long nrows = ln.ncols * ln.nplanes;
        auto kfn = ln.in_tw ? k_dst_half_rows<Ti, T, To, M, TPB, true> : k_dst_half_rows<Ti, T, To, M, TPB, false>;
        if (shm > 65536) {
            hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            if (e != hipSuccess) return e;
        }
        long grid = (nrows + cfg::LPB - 1) / cfg::LPB;
        const RowMap rin{ln.N, ln.ny, ln.in_tw, ln.field, 0.0}, rout{ln.N, ln.ny, ln.out_tw, ln.field, ln.acc};
        hipLaunchKernelGGL(kfn, dim3((unsigned)grid), dim3(cfg::THREADS), shm, st, pi, po, nrows, scale, pa, rin,
                           rout);
        return hipGetLastError();
    }
    if (ln.col_stride != 1 || ln.acc != 0) return hipErrorInvalidValue;   // strided lines: plain stores
    if constexpr (TPB == 256 && M >= 128 && M <= 256) {
        using wcfg = half_cfg<M, 512>;
        constexpr int WNW = wcfg::P > 64 ? wcfg::P / 64 : 1;
        constexpr size_t wshm = sizeof(T) * (wcfg::LPB * wcfg::LS + 2 * wcfg::LPB * WNW);
        if (wide && wshm <= 160 * 1024) {   // tiles of 512/P lines: twice the row-segment width
            auto kfn = k_dst_half_cols<Ti, T, To, M, 512>;
            if (wshm > 65536) {
                hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)wshm);
                if (e != hipSuccess) return e;
            }
            long ntiles = (ln.ncols + wcfg::LPB - 1) / wcfg::LPB;
            long grid = ntiles * ln.nplanes;
            hipLaunchKernelGGL(kfn, dim3((unsigned)grid), dim3(wcfg::THREADS), wshm, st, pi, po, ln.ncols, ntiles,
                               ln.plane_stride, ln.elem_stride, scale, pa);
            return hipGetLastError();
        }
    }
    auto kfn = k_dst_half_cols<Ti, T, To, M, TPB>;
    if (shm > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    long ntiles = (ln.ncols + cfg::LPB - 1) / cfg::LPB;
    long grid = ntiles * ln.nplanes;
    hipLaunchKernelGGL(kfn, dim3((unsigned)grid), dim3(cfg::THREADS), shm, st, pi, po, ln.ncols, ntiles,
                       ln.plane_stride, ln.elem_stride, scale, pa);
    return hipGetLastError();
}

template <typename Ti, typename T, typename To, int TPB>
static hipError_t launch_fft_t(const Lines &ln, int algo, const void *in, void *out, const void *tw, hipStream_t st) {
    const bool wide = (algo & 4) != 0;
    if (!(algo & 1)) {
        switch (ln.N + 1) {
        case 2: return launch_half_M<Ti, T, To, 2, TPB>(ln, in, out, tw, wide, st);
        case 4: return launch_half_M<Ti, T, To, 4, TPB>(ln, in, out, tw, wide, st);
        case 8: return launch_half_M<Ti, T, To, 8, TPB>(ln, in, out, tw, wide, st);
        case 16: return launch_half_M<Ti, T, To, 16, TPB>(ln, in, out, tw, wide, st);
        case 32: return launch_half_M<Ti, T, To, 32, TPB>(ln, in, out, tw, wide, st);
        case 64: return launch_half_M<Ti, T, To, 64, TPB>(ln, in, out, tw, wide, st);
        case 128: return launch_half_M<Ti, T, To, 128, TPB>(ln, in, out, tw, wide, st);
        case 256: return launch_half_M<Ti, T, To, 256, TPB>(ln, in, out, tw, wide, st);
        case 512: return launch_half_M<Ti, T, To, 512, TPB>(ln, in, out, tw, wide, st);
        case 1024: return launch_half_M<Ti, T, To, 1024, TPB>(ln, in, out, tw, wide, st);
        case 2048: return launch_half_M<Ti, T, To, 2048, TPB>(ln, in, out, tw, wide, st);
        case 4096: return launch_half_M<Ti, T, To, 4096, TPB>(ln, in, out, tw, wide, st);
        default: return hipErrorInvalidValue;
        }
    }
    switch (2 * (ln.N + 1)) {
    case 4: return launch_fft_L<Ti, T, To, 4, TPB>(ln, in, out, tw, st);
    case 8: return launch_fft_L<Ti, T, To, 8, TPB>(ln, in, out, tw, st);
    case 16: return launch_fft_L<Ti, T, To, 16, TPB>(ln, in, out, tw, st);
    case 32: return launch_fft_L<Ti, T, To, 32, TPB>(ln, in, out, tw, st);
    case 64: return launch_fft_L<Ti, T, To, 64, TPB>(ln, in, out, tw, st);
    case 128: return launch_fft_L<Ti, T, To, 128, TPB>(ln, in, out, tw, st);
    case 256: return launch_fft_L<Ti, T, To, 256, TPB>(ln, in, out, tw, st);
    case 512: return launch_fft_L<Ti, T, To, 512, TPB>(ln, in, out, tw, st);
    case 1024: return launch_fft_L<Ti, T, To, 1024, TPB>(ln, in, out, tw, st);
    case 2048: return launch_fft_L<Ti, T, To, 2048, TPB>(ln, in, out, tw, st);
    case 4096: return launch_fft_L<Ti, T, To, 4096, TPB>(ln, in, out, tw, st);
    case 8192: return launch_fft_L<Ti, T, To, 8192, TPB>(ln, in, out, tw, st);
    default: return hipErrorInvalidValue;
    }
}

// algo bit 0: odd-extension transform (else half length); bit 1: 64-thread
// blocks (else 256); bit 2: 512-thread blocks for strided lines of the
// half-length transform with 128 <= N+1 <= 256 (tiles of twice the lines)
template <typename Ti, typename T, typename To>
static hipError_t launch_dst_t(const Lines &ln, int algo, const void *in, void *out, const void *tw, const void *S,
                               hipStream_t st) {
    if (fft_supported(ln.N)) {
        if (algo & 2) return launch_fft_t<Ti, T, To, 64>(ln, algo, in, out, tw, st);
        return launch_fft_t<Ti, T, To, 256>(ln, algo, in, out, tw, st);
    }
    if (in == out || S == nullptr) return hipErrorInvalidValue;
    long nlines = ln.ncols * ln.nplanes;
    dim3 grid((unsigned)((nlines + 31) / 32), (unsigned)((ln.N + 63) / 64));
    if (algo & 8)   // vector-FMA form (kept for comparison)
        hipLaunchKernelGGL((k_dst_gemm<Ti, T, To>), grid, dim3(256), 0, st, (const C2<Ti> *)in, (C2<To> *)out, ln,
                           nlines, (const T *)S);
    else
        hipLaunchKernelGGL((k_dst_mfma<Ti, T, To>), grid, dim3(256), 0, st, (const C2<Ti> *)in, (C2<To> *)out, ln,
                           nlines, (const T *)S);
    return hipGetLastError();
}

hipError_t launch_dst(const Lines &ln, DstTypes ty, const void *in, void *out, const void *tw, const void *S,
                      hipStream_t st) {
    if (ln.N < 1) return hipErrorInvalidValue;
    const int a = ty.algo;
    if (ty.comp == 0) {
        if (ty.in == 0 && ty.out == 0) return launch_dst_t<double, double, double>(ln, a, in, out, tw, S, st);
        return hipErrorInvalidValue;
    }
    if (ty.in == 0 && ty.out == 1) return launch_dst_t<double, float, float>(ln, a, in, out, tw, S, st);
    if (ty.in == 1 && ty.out == 1) return launch_dst_t<float, float, float>(ln, a, in, out, tw, S, st);
    if (ty.in == 1 && ty.out == 0) return launch_dst_t<float, float, double>(ln, a, in, out, tw, S, st);
    return hipErrorInvalidValue;
}

}  // namespace hh
