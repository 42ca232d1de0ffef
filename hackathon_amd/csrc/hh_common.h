// Shared device/host definitions for the MI355X compact-Helmholtz library.
// Complex values are interleaved (re, im) pairs: double2 for FP64, float2
// for FP32 -- the same bytes as C99 `double _Complex` used by the host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HH_DEV __device__ __forceinline__

template <typename T> struct cvec;
template <> struct cvec<double> { using t = double2; };
template <> struct cvec<float> { using t = float2; };
template <typename T> using C2 = typename cvec<T>::t;

template <typename T> HH_DEV C2<T> mk(T r, T i) { C2<T> z; z.x = r; z.y = i; return z; }
template <typename T> HH_DEV C2<T> cadd(C2<T> a, C2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> HH_DEV C2<T> csub(C2<T> a, C2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> HH_DEV C2<T> cneg(C2<T> a) { return mk<T>(-a.x, -a.y); }
template <typename T> HH_DEV C2<T> cmul(C2<T> a, C2<T> b) {
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> HH_DEV C2<T> cscale(C2<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// a*b + c
template <typename T> HH_DEV C2<T> cfma(C2<T> a, C2<T> b, C2<T> c) {
    return mk<T>(c.x + a.x * b.x - a.y * b.y, c.y + a.x * b.y + a.y * b.x);
}
// c - a*b
template <typename T> HH_DEV C2<T> cfms(C2<T> a, C2<T> b, C2<T> c) {
    return mk<T>(c.x - (a.x * b.x - a.y * b.y), c.y - (a.x * b.y + a.y * b.x));
}
// multiply by -i
template <typename T> HH_DEV C2<T> cmul_mi(C2<T> a) { return mk<T>(a.y, -a.x); }
// multiply by +i
template <typename T> HH_DEV C2<T> cmul_i(C2<T> a) { return mk<T>(-a.y, a.x); }
// Reciprocal of a positive real: hardware estimate (v_rcp_f64 / v_rcp_f32)
// refined by Newton steps in residual form, e = 1 - n r, r += r e (fused).
// Each step squares the relative error; two steps take an estimate good to
// 2^-14 or better to the last bit of FP64, one step suffices for FP32.  The
// quotient is not the correctly rounded one but stays within an ulp of it.
HH_DEV double rcp_nr(double n) {
    double r = __builtin_amdgcn_rcp(n);
    double e = __builtin_fma(-n, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-n, r, 1.0);
    return __builtin_fma(r, e, r);
}
HH_DEV float rcp_nr(float n) {
    float r = __builtin_amdgcn_rcpf(n);
    float e = __builtin_fmaf(-n, r, 1.0f);
    return __builtin_fmaf(r, e, r);
}
// 1/z = conj(z)/|z|^2
template <typename T> HH_DEV C2<T> crcp(C2<T> z) {
    T n = z.x * z.x + z.y * z.y;
    T r = rcp_nr(n);
    return mk<T>(z.x * r, -z.y * r);
}
template <typename T> HH_DEV T cabs2(C2<T> z) { return z.x * z.x + z.y * z.y; }

template <typename To, typename Ti> HH_DEV C2<To> cconv(C2<Ti> z) { return mk<To>((To)z.x, (To)z.y); }

// Bounds of the kernels: maximum system length handled by one tridiagonal
// block and maximum FFT length.
#define HH_FFT_LMAX 8192
