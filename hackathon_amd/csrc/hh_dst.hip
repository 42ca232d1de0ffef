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

template <int L> struct dst_cfg {
    static constexpr int P = fft_shape<L>::P;
    static constexpr int LPB = P >= 256 ? 1 : 256 / P;   // lines per block
    static constexpr int THREADS = P * LPB;
    static constexpr int LP = fft_shape<L>::LP;
};

template <typename Ti, typename T, typename To, int L>
__global__ __launch_bounds__(dst_cfg<L>::THREADS) void k_dst_fft_rows(
    const C2<Ti> *in, C2<To> *out, long nrows, T scale,
    const C2<T> *__restrict__ tw) {
    constexpr int P = dst_cfg<L>::P, LPB = dst_cfg<L>::LPB, LP = dst_cfg<L>::LP;
    constexpr int M = L / 2, N = M - 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    const int li = threadIdx.x / P, t = threadIdx.x % P;
    T *sre = smem + (2 * li) * LP, *sim = smem + (2 * li + 1) * LP;
    const long row = (long)blockIdx.x * LPB + li;
    const bool valid = row < nrows;
    const C2<Ti> *x = in + (valid ? row : 0) * N;
    C2<To> *y = out + (valid ? row : 0) * N;
    auto get = [&](int n) -> C2<T> {
        if (!valid || n == 0 || n == M) return mk<T>(0, 0);
        if (n < M) return cconv<T, Ti>(x[n - 1]);
        return cneg<T>(cconv<T, Ti>(x[L - n - 1]));
    };
    auto put = [&](int k, C2<T> z) {
        if (valid && k >= 1 && k <= N) y[k - 1] = cconv<To, T>(mk<T>(-z.y * scale, z.x * scale));
    };
    fft_line<T, L, false>(sre, sim, tw, t, get, put);
}

template <typename Ti, typename T, typename To, int L>
__global__ __launch_bounds__(dst_cfg<L>::THREADS) void k_dst_fft_cols(
    const C2<Ti> *in, C2<To> *out, long ncols, long ntiles,
    long plane_stride, long elem_stride, T scale, const C2<T> *__restrict__ tw) {
    constexpr int P = dst_cfg<L>::P, LPB = dst_cfg<L>::LPB, LP = dst_cfg<L>::LP;
    constexpr int M = L / 2, N = M - 1, NT = dst_cfg<L>::THREADS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = reinterpret_cast<T *>(smem_raw);
    const long tile = blockIdx.x % ntiles, plane = blockIdx.x / ntiles;
    const long c0 = tile * LPB;
    const long base = plane * plane_stride + c0;
    // coalesced tile load: lanes run along the columns of one row
    for (int e = threadIdx.x; e < N * LPB; e += NT) {
        int c = e % LPB, q = e / LPB;
        T *lre = smem + (2 * c) * LP, *lim = smem + (2 * c + 1) * LP;
        C2<T> v = mk<T>(0, 0);
        if (c0 + c < ncols) v = cconv<T, Ti>(in[base + c + q * elem_stride]);
        lre[lds_pad(q + 1)] = v.x;
        lim[lds_pad(q + 1)] = v.y;
    }
    __syncthreads();
    const int li = threadIdx.x / P, t = threadIdx.x % P;
    T *sre = smem + (2 * li) * LP, *sim = smem + (2 * li + 1) * LP;
    auto get = [&](int n) -> C2<T> {
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
    for (int e = threadIdx.x; e < N * LPB; e += NT) {
        int c = e % LPB, q = e / LPB;
        if (c0 + c < ncols) {
            const T *lre = smem + (2 * c) * LP, *lim = smem + (2 * c + 1) * LP;
            out[base + c + q * elem_stride] = cconv<To, T>(mk<T>(lre[lds_pad(q + 1)], lim[lds_pad(q + 1)]));
        }
    }
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
            if (k < N) out[bse + (long)k * ln.elem_stride] = cconv<To, T>(acc[u]);
        }
    }
}

bool fft_supported(int N) {
    int L = 2 * (N + 1);
    return N >= 1 && L <= HH_FFT_LMAX && (L & (L - 1)) == 0 && L >= 4;
}

template <typename Ti, typename T, typename To, int L>
static hipError_t launch_fft_L(const Lines &ln, const void *in, void *out, const void *tw, hipStream_t st) {
    using cfg = dst_cfg<L>;
    size_t shm = sizeof(T) * 2 * cfg::LPB * cfg::LP;
    T scale = (T)sqrt(0.5 / (double)(L / 2));
    const C2<Ti> *pi = (const C2<Ti> *)in;
    C2<To> *po = (C2<To> *)out;
    const C2<T> *ptw = (const C2<T> *)tw;
    if (ln.elem_stride == 1 && ln.col_stride == ln.N && ln.nplanes * ln.ncols > 0 &&
        (ln.nplanes == 1 || ln.plane_stride == ln.ncols * ln.N)) {
        long nrows = ln.ncols * ln.nplanes;
        auto kfn = k_dst_fft_rows<Ti, T, To, L>;
        if (shm > 65536) {
            hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            if (e != hipSuccess) return e;
        }
        long grid = (nrows + cfg::LPB - 1) / cfg::LPB;
        hipLaunchKernelGGL(kfn, dim3((unsigned)grid), dim3(cfg::THREADS), shm, st, pi, po, nrows, scale, ptw);
        return hipGetLastError();
    }
    if (ln.col_stride != 1) return hipErrorInvalidValue;
    auto kfn = k_dst_fft_cols<Ti, T, To, L>;
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

template <typename Ti, typename T, typename To>
static hipError_t launch_dst_t(const Lines &ln, const void *in, void *out, const void *tw, const void *S,
                               hipStream_t st) {
    if (fft_supported(ln.N)) {
        switch (2 * (ln.N + 1)) {
        case 4: return launch_fft_L<Ti, T, To, 4>(ln, in, out, tw, st);
        case 8: return launch_fft_L<Ti, T, To, 8>(ln, in, out, tw, st);
        case 16: return launch_fft_L<Ti, T, To, 16>(ln, in, out, tw, st);
        case 32: return launch_fft_L<Ti, T, To, 32>(ln, in, out, tw, st);
        case 64: return launch_fft_L<Ti, T, To, 64>(ln, in, out, tw, st);
        case 128: return launch_fft_L<Ti, T, To, 128>(ln, in, out, tw, st);
        case 256: return launch_fft_L<Ti, T, To, 256>(ln, in, out, tw, st);
        case 512: return launch_fft_L<Ti, T, To, 512>(ln, in, out, tw, st);
        case 1024: return launch_fft_L<Ti, T, To, 1024>(ln, in, out, tw, st);
        case 2048: return launch_fft_L<Ti, T, To, 2048>(ln, in, out, tw, st);
        case 4096: return launch_fft_L<Ti, T, To, 4096>(ln, in, out, tw, st);
        case 8192: return launch_fft_L<Ti, T, To, 8192>(ln, in, out, tw, st);
        default: return hipErrorInvalidValue;
        }
    }
    if (in == out || S == nullptr) return hipErrorInvalidValue;
    long nlines = ln.ncols * ln.nplanes;
    dim3 grid((unsigned)((nlines + 31) / 32), (unsigned)((ln.N + 63) / 64));
    hipLaunchKernelGGL((k_dst_gemm<Ti, T, To>), grid, dim3(256), 0, st, (const C2<Ti> *)in, (C2<To> *)out, ln,
                       nlines, (const T *)S);
    return hipGetLastError();
}

hipError_t launch_dst(const Lines &ln, DstTypes ty, const void *in, void *out, const void *tw, const void *S,
                      hipStream_t st) {
    if (ln.N < 1) return hipErrorInvalidValue;
    if (ty.comp == 0) {
        if (ty.in == 0 && ty.out == 0) return launch_dst_t<double, double, double>(ln, in, out, tw, S, st);
        return hipErrorInvalidValue;
    }
    if (ty.in == 0 && ty.out == 1) return launch_dst_t<double, float, float>(ln, in, out, tw, S, st);
    if (ty.in == 1 && ty.out == 1) return launch_dst_t<float, float, float>(ln, in, out, tw, S, st);
    if (ty.in == 1 && ty.out == 0) return launch_dst_t<float, float, double>(ln, in, out, tw, S, st);
    return hipErrorInvalidValue;
}

}  // namespace hh
