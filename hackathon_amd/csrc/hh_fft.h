// Device-side Stockham FFT used by the DST-I kernels.
//
// A line of length L = 2^p is transformed by P = L/V threads, each holding
// V = min(16, L) complex values in registers.  Stages are radix-16 (the last
// one radix L/Ns when fewer points remain); between stages the values are
// exchanged through LDS (split real / imaginary arrays, padded by one
// element per 16 to break power-of-two bank strides; padding per 32 removes
// more two-way conflicts on paper but measured 1-2% slower).  Stage s with radix R
// and span Ns (Stockham autosort, decimation in time):
//   butterfly j in [0, L/R):  k = j mod Ns
//     in  x[r] = buf[j + r L/R]                 r = 0..R-1
//     x[r] *= W_{Ns R}^{k r}                    (table tw[q] = e^{-2 pi i q/L})
//     x = DFT_R(x)
//     out buf[(j/Ns) Ns R + k + r Ns] = x[r]
// After the last stage the outputs are in natural order; thread t then
// holds Z[j + r L/R] for its butterflies j = t + b P.
#pragma once
#include "hh_common.h"

namespace hh {

HH_DEV int lds_pad(int i) { return i + (i >> 4); }

template <typename T> HH_DEV void dft2(C2<T>* v) {
    C2<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
}

template <typename T> HH_DEV void dft4(C2<T>& v0, C2<T>& v1, C2<T>& v2, C2<T>& v3) {
    C2<T> t0 = cadd<T>(v0, v2), t1 = csub<T>(v0, v2);
    C2<T> t2 = cadd<T>(v1, v3), t3 = cmul_mi<T>(csub<T>(v1, v3));
    v0 = cadd<T>(t0, t2);
    v2 = csub<T>(t0, t2);
    v1 = cadd<T>(t1, t3);
    v3 = csub<T>(t1, t3);
}

template <typename T> HH_DEV void dft4(C2<T>* v) { dft4<T>(v[0], v[1], v[2], v[3]); }

// W8^1 * a and W8^3 * a
template <typename T> HH_DEV C2<T> mul_w8_1(C2<T> a) {
    const T c = (T)0.70710678118654752440;
    return mk<T>(c * (a.x + a.y), c * (a.y - a.x));
}
template <typename T> HH_DEV C2<T> mul_w8_3(C2<T> a) {
    const T c = (T)0.70710678118654752440;
    return mk<T>(c * (a.y - a.x), -c * (a.x + a.y));
}

template <typename T> HH_DEV void dft8(C2<T>* v) {
    C2<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C2<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<T>(e0, e1, e2, e3);
    dft4<T>(o0, o1, o2, o3);
    o1 = mul_w8_1<T>(o1);
    o2 = cmul_mi<T>(o2);
    o3 = mul_w8_3<T>(o3);
    v[0] = cadd<T>(e0, o0); v[4] = csub<T>(e0, o0);
    v[1] = cadd<T>(e1, o1); v[5] = csub<T>(e1, o1);
    v[2] = cadd<T>(e2, o2); v[6] = csub<T>(e2, o2);
    v[3] = cadd<T>(e3, o3); v[7] = csub<T>(e3, o3);
}

// 16 = 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2
template <typename T> HH_DEV void dft16(C2<T>* v) {
    const T c1 = (T)0.92387953251128675613;  // cos(pi/8)
    const T s1 = (T)0.38268343236508977173;  // sin(pi/8)
    C2<T> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        a[n2][0] = v[n2]; a[n2][1] = v[4 + n2]; a[n2][2] = v[8 + n2]; a[n2][3] = v[12 + n2];
        dft4<T>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
    }
    // twiddles W16^{n2 k1}
    a[1][1] = cmul<T>(a[1][1], mk<T>(c1, -s1));
    a[1][2] = mul_w8_1<T>(a[1][2]);
    a[1][3] = cmul<T>(a[1][3], mk<T>(s1, -c1));
    a[2][1] = mul_w8_1<T>(a[2][1]);
    a[2][2] = cmul_mi<T>(a[2][2]);
    a[2][3] = mul_w8_3<T>(a[2][3]);
    a[3][1] = cmul<T>(a[3][1], mk<T>(s1, -c1));
    a[3][2] = mul_w8_3<T>(a[3][2]);
    a[3][3] = cmul<T>(a[3][3], mk<T>(-c1, s1));
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        C2<T> b0 = a[0][k1], b1 = a[1][k1], b2 = a[2][k1], b3 = a[3][k1];
        dft4<T>(b0, b1, b2, b3);
        v[k1] = b0; v[k1 + 4] = b1; v[k1 + 8] = b2; v[k1 + 12] = b3;
    }
}

template <typename T, int R> HH_DEV void dftR(C2<T>* v) {
    if constexpr (R == 2) dft2<T>(v);
    else if constexpr (R == 4) dft4<T>(v);
    else if constexpr (R == 8) dft8<T>(v);
    else if constexpr (R == 16) dft16<T>(v);
}

// Twiddles w[r] = W_{NS R}^{k r}, r < R, of butterfly k from the table
// tw[q] = e^{-2 pi i q/L}.  Radix 8 and 16 split r = r1 + 4 r2: three loads
// of W^{k r1}, R/4 - 1 of W^{4 k r2}, and one complex product for each of
// the other twiddles (one rounding more than a table entry).
template <typename T, int L, int R, int NS>
HH_DEV void stage_twiddles(const C2<T>* __restrict__ tw, int k, C2<T>* w) {
    constexpr int S = L / (NS * R);
    w[0] = mk<T>(1, 0);
    if constexpr (R < 8) {
#pragma unroll
        for (int r = 1; r < R; r++) w[r] = tw[k * r * S];
    } else {
        C2<T> lo[4], hi[R / 4];
#pragma unroll
        for (int r1 = 1; r1 < 4; r1++) lo[r1] = tw[k * r1 * S];
#pragma unroll
        for (int r2 = 1; r2 < R / 4; r2++) hi[r2] = tw[k * 4 * r2 * S];
#pragma unroll
        for (int r = 1; r < R; r++) {
            const int r1 = r % 4, r2 = r / 4;
            w[r] = r2 == 0 ? lo[r1] : r1 == 0 ? hi[r2] : cmul<T>(hi[r2], lo[r1]);
        }
    }
}

template <int L> struct fft_shape {
    static constexpr int V = L < 16 ? L : 16;   // values per thread
    static constexpr int P = L / V;             // threads per line
    static constexpr int LP = L + L / 16 + 1;   // padded LDS line length
};

// One Stockham stage.  FIRST: inputs come from get(n, r), r = n / P the
// value's slot (compile time; the first stage has radix V, so n = t + r P);
// otherwise from LDS.
// LAST: outputs go to put(k, z) after an optional barrier (set LASTBAR when
// put() writes to the LDS line buffer); otherwise they go back to LDS.
template <typename T, int L, int R, int NS, bool FIRST, bool LAST, bool LASTBAR, class Get, class Put>
HH_DEV void fft_stage(T* sre, T* sim, const C2<T>* __restrict__ tw, int t, Get& get, Put& put) {
    constexpr int V = fft_shape<L>::V;
    constexpr int P = fft_shape<L>::P;
    constexpr int NB = V / R;
    C2<T> v[V];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        int j = t + b * P;
#pragma unroll
        for (int r = 0; r < R; r++) {
            int idx = j + r * (L / R);
            if constexpr (FIRST) v[b * R + r] = get(idx, r);
            else v[b * R + r] = mk<T>(sre[lds_pad(idx)], sim[lds_pad(idx)]);
        }
    }
    // L, NS and P are powers of two: either NS divides P, and every
    // butterfly j = t + b P of the thread has k = t mod NS, or the thread's
    // butterflies all lie below NS (NB P = NS), k = t + b P and
    // W^{k r} = W^{t r} W^{b P r}, the second factor the same in every lane
    // (a wave-uniform load).  Either way the per-lane twiddles are loaded
    // once per stage, not once per butterfly.
    static_assert(NS <= P || NB * P == NS, "unexpected stage shape");
    C2<T> w[R];
    if constexpr (NS > 1) stage_twiddles<T, L, R, NS>(tw, NS <= P ? (t & (NS - 1)) : t, w);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                C2<T> x = w[r];
                if constexpr (NS > P) {
                    if (b > 0) x = cmul<T>(x, tw[b * P * r * (L / (NS * R))]);
                }
                v[b * R + r] = cmul<T>(v[b * R + r], x);
            }
        }
        dftR<T, R>(&v[b * R]);
    }
    if constexpr (!LAST) {
        __syncthreads();
#pragma unroll
        for (int b = 0; b < NB; b++) {
            int j = t + b * P;
            int k = j % NS;
            int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) {
                int o = lds_pad(base + r * NS);
                sre[o] = v[b * R + r].x;
                sim[o] = v[b * R + r].y;
            }
        }
        __syncthreads();
    } else {
        if constexpr (LASTBAR) __syncthreads();
#pragma unroll
        for (int b = 0; b < NB; b++) {
            int j = t + b * P;
#pragma unroll
            for (int r = 0; r < R; r++) put(j + r * NS, v[b * R + r]);
        }
    }
}

template <typename T, int L, int NS, bool LASTBAR, class Get, class Put>
HH_DEV void fft_run(T* sre, T* sim, const C2<T>* __restrict__ tw, int t, Get& get, Put& put) {
    constexpr int V = fft_shape<L>::V;
    constexpr int R = (L / NS) >= V ? V : (L / NS);
    constexpr bool FIRST = (NS == 1);
    constexpr bool LAST = (NS * R == L);
    fft_stage<T, L, R, NS, FIRST, LAST, LASTBAR>(sre, sim, tw, t, get, put);
    if constexpr (!LAST) fft_run<T, L, NS * R, LASTBAR>(sre, sim, tw, t, get, put);
}

// Full FFT of one line: get(n) supplies z_n (n in [0,L)), put(k, Z_k) receives
// the transform.  All threads of the block must call it (it has barriers).
template <typename T, int L, bool LASTBAR, class Get, class Put>
HH_DEV void fft_line(T* sre, T* sim, const C2<T>* __restrict__ tw, int t, Get& get, Put& put) {
    fft_run<T, L, 1, LASTBAR>(sre, sim, tw, t, get, put);
}

}  // namespace hh
