// C-ABI layer of the MI355X compact-Helmholtz library (include/hh_solver.h).
//
// Contexts own the stream, the per-length twiddle tables, the dense DST
// matrices of non power-of-two sizes and a growable scratch buffer.  Plans
// own the device copies of the O(N) stencil coefficient vectors and the
// eigenvalues 2cos(k pi/(N+1)); they replace the reference's LU() factors.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hh_solver.h"
#include "hh_internal.h"

using cd = std::complex<double>;

static thread_local std::string g_err;

// Workspace of hh_solve_noalloc, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
// one array per device, held until the process ends.  In host code the
// array's name is a host-side stand-in: its device address is looked up once
// per device at context creation (hipGetSymbolAddress) and only that address
// is ever passed to kernels.
#define HH_NOALLOC_BYTES ((size_t)128 << 20)
static __device__ __align__(256) unsigned char noalloc_scratch[HH_NOALLOC_BYTES];
static std::mutex g_noalloc_mu;
static std::map<int, std::pair<void *, hipError_t>> g_noalloc;   // device -> (address, lookup status)

static int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(HH_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                        __FILE__, __LINE__);                                                       \
    } while (0)

struct hh_context {
    int device = 0;
    hipStream_t own = nullptr, stream = nullptr;
    std::map<int, void *> tw64, tw32;   // twiddles by FFT length
    std::map<int, void *> S64, S32;     // dense DST matrices by N
    std::map<int, void *> th64, th32;   // half-length transform tables by N+1
    int dst_algo = -1;                  // -1 auto; bit 0: odd extension (else (N+1)-point), bit 1: 64-thread blocks
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    double *d_norm = nullptr;
    size_t budget = 0, in_use = 0;
    void *noalloc_ws = nullptr;         // this device's noalloc_scratch (device address)
    hipError_t noalloc_status = hipSuccess;
    bool no_alloc = false;              // inside hh_solve_noalloc: allocations are refused
    std::mutex mu;
    // kernel timing (hh_context_enable_timing)
    struct Rec { hipEvent_t a, b; int cat; double bytes; };
    bool timing = false;
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double acc_ms[5] = {0, 0, 0, 0, 0}, acc_bytes[5] = {0, 0, 0, 0, 0};
    long acc_n[5] = {0, 0, 0, 0, 0};
};

// Records a pair of events around the launches in its scope when timing is on.
// Phase ranges for rocprofv3 --marker-trace (the reference's NVTX ranges,
// 2D_Program/solver_functions/cuda_helper.h:13-25 around the phases of
// solver.c:16-92): one roctx range per kernel pass, named by category, and
// one per solve; they cost a call each when no profiler listens.
static const char *const kPhaseName[] = {"hh:dst", "hh:dst_dense", "hh:tridiagonal", "hh:residual_update",
                                         "hh:rhs"};
struct Range {
    explicit Range(const char *name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
};

struct TimeScope {
    hh_context *c;
    hh_context::Rec r;
    bool on;
    Range range;
    TimeScope(hh_context *ctx, int cat, double bytes)
        : c(ctx), on(ctx->timing && !ctx->no_alloc), range(kPhaseName[cat]) {
        if (!on) return;
        r.cat = cat;
        r.bytes = bytes;
        r.a = take();
        r.b = take();
        (void)hipEventRecord(r.a, c->stream);
    }
    hipEvent_t take() {
        if (!c->pool.empty()) { hipEvent_t e = c->pool.back(); c->pool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    ~TimeScope() {
        if (!on) return;
        (void)hipEventRecord(r.b, c->stream);
        c->recs.push_back(r);
    }
};

struct hh_plan {
    hh_context *ctx = nullptr;
    int dim = 2, Nx = 0, Ny = 0, Nz = 1, prec = HH_PREC_F64;
    int nop = 1;                                 // operators (one per right-hand side b % nop)
    void *coef64 = nullptr, *coef32 = nullptr;   // [op][q][6] or [op][q][12]
    double *lamx = nullptr, *lamy = nullptr;
    std::vector<cd> hcoef;                       // host copy of the band table
    std::vector<float> h32;                      // FP32 copy staged for hh_plan_set_operators
    // real parts of the table, same order (read as (re c_2i, re c_2i+1) pairs by
    // the real-band tridiagonal kernels), FP64 and FP32
    void *coefr64 = nullptr, *coefr32 = nullptr;
    std::vector<double> hr64;
    std::vector<float> hr32;
    int real_op = 0;                             // every coefficient of the current operators real
    int real_mode = -1;                          // -1: real-band kernels when real_op; 0: complex-band always
    int cap_nop = 1;                             // operator slots allocated on the device
    double tol = 1e-12;
    int max_iter = 6;
    int tri_W = 0, tri_mm = 0;                   // tuned tridiagonal block shape (0 = heuristic)
    int tri_nr = 1;                              // right-hand sides per thread of the tuned shape
    // 2D spectrum layout: 0 row-major (default: measured fastest for the whole
    // solve, 1753 us against 1761-1844 us for tiles of 4..32 modes at N=1023,
    // 64 right-hand sides), -1 tiles of the block width, else the tile width
    int tile_w = 0;
    int last_sweeps = 0;
    double last_rel = 0;
    // eigenvalues 2cos(k pi/(N+1)) (host copies of lamx / lamy)
    std::vector<double> hlamx, hlamy;
    // near-resonant modes of the current operators and their reference-order
    // factors (hh_modes.hip), with their device copies
    double kappa_flag = hh::HH_KAPPA_FLAG;
    hh::ModeSet ms;
    int *d_foff = nullptr, *d_fmode = nullptr;
    uint32_t *d_bits = nullptr;
    void *d_facc = nullptr, *d_facr = nullptr;
    std::vector<cd> hfacc_t;                     // staging of the device layout
    std::vector<double> hfacr_t;
    size_t cap_foff = 0, cap_fmode = 0, cap_bits = 0, cap_facc = 0, cap_facr = 0;
    // Generation: bumped whenever a change would invalidate a recorded graph
    // (device tables reallocated, band kind, operator count or launch shapes
    // changed); graphs record it at capture and refuse to replay a newer plan.
    long gen = 0;
};

// ------------------------------------------------------------------ helpers
static int dev_alloc(hh_context *c, void **p, size_t bytes) {
    if (c->no_alloc)
        return fail(HH_ERR_UNSUPPORTED, "hh_solve_noalloc: a table of %zu bytes is missing (create the plan with the "
                                        "context's current DST algorithm, or solve once with hh_solve first)", bytes);
    if (c->budget && c->in_use + bytes > c->budget)
        return fail(HH_ERR_NOMEM, "allocation of %zu bytes exceeds the memory budget (%zu of %zu in use)", bytes,
                    c->in_use, c->budget);
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(HH_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    c->in_use += bytes;
    return HH_OK;
}

static void dev_free(hh_context *c, void *p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    c->in_use -= bytes < c->in_use ? bytes : c->in_use;
}

static int ensure_scratch(hh_context *c, size_t bytes) {
    if (c->scratch_bytes >= bytes) return HH_OK;
    if (c->scratch) {
        HIPCHK(hipStreamSynchronize(c->stream));
        dev_free(c, c->scratch, c->scratch_bytes);
        c->scratch = nullptr;
        c->scratch_bytes = 0;
    }
    int rc = dev_alloc(c, &c->scratch, bytes);
    if (rc) return rc;
    c->scratch_bytes = bytes;
    return HH_OK;
}

static int set_device(hh_context *c) {
    HIPCHK(hipSetDevice(c->device));
    return HH_OK;
}

// e^{-2 pi i q/L}, q < L, with exact argument reduction to the first octant
static void twiddles(int L, std::vector<cd> &out) {
    out.resize(L);
    for (int q = 0; q < L; q++) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)q / (long double)L;
        out[q] = cd((double)cosl(a), (double)sinl(a));
    }
}

static int get_twiddles(hh_context *c, int N, int f32, void **out) {
    int L = 2 * (N + 1);
    auto &m = f32 ? c->tw32 : c->tw64;
    auto it = m.find(L);
    if (it != m.end()) { *out = it->second; return HH_OK; }
    std::vector<cd> tw;
    twiddles(L, tw);
    void *d = nullptr;
    size_t bytes = (size_t)L * (f32 ? 8 : 16);
    int rc = dev_alloc(c, &d, bytes);
    if (rc) return rc;
    if (f32) {
        std::vector<float> h(2 * L);
        for (int q = 0; q < L; q++) { h[2 * q] = (float)tw[q].real(); h[2 * q + 1] = (float)tw[q].imag(); }
        HIPCHK(hipMemcpy(d, h.data(), bytes, hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemcpy(d, tw.data(), bytes, hipMemcpyHostToDevice));
    }
    m[L] = d;
    *out = d;
    return HH_OK;
}

// S[n][k] = sqrt(2/(N+1)) sin(pi (n+1)(k+1)/(N+1)), exact integer reduction
static int get_dst_matrix(hh_context *c, int N, int f32, void **out) {
    auto &m = f32 ? c->S32 : c->S64;
    auto it = m.find(N);
    if (it != m.end()) { *out = it->second; return HH_OK; }
    std::vector<double> S((size_t)N * N);
    long M2 = 2L * (N + 1);
    double s = std::sqrt(2.0 / (N + 1));
    for (long n = 0; n < N; n++)
        for (long k = 0; k < N; k++) {
            long q = ((n + 1) * (k + 1)) % M2;
            S[n * N + k] = s * (double)sinl(3.14159265358979323846264338327950288L * (long double)q / (long double)(N + 1));
        }
    void *d = nullptr;
    size_t bytes = (size_t)N * N * (f32 ? 4 : 8);
    int rc = dev_alloc(c, &d, bytes);
    if (rc) return rc;
    if (f32) {
        std::vector<float> h(S.begin(), S.end());
        HIPCHK(hipMemcpy(d, h.data(), bytes, hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemcpy(d, S.data(), bytes, hipMemcpyHostToDevice));
    }
    m[N] = d;
    *out = d;
    return HH_OK;
}

// Table of the half-length transform: M = N+1 twiddles e^{-2 pi i q/M}, then
// the M reals sin(pi j/M) (packed into (M+1)/2 complex slots).
static int get_half_table(hh_context *c, int N, int f32, void **out) {
    const int M = N + 1;
    auto &m = f32 ? c->th32 : c->th64;
    auto it = m.find(M);
    if (it != m.end()) { *out = it->second; return HH_OK; }
    std::vector<cd> tw;
    twiddles(M, tw);
    const int ns = (M + 1) / 2;
    std::vector<double> h(2 * ((size_t)M + ns), 0.0);
    for (int q = 0; q < M; q++) { h[2 * q] = tw[q].real(); h[2 * q + 1] = tw[q].imag(); }
    for (int j = 0; j < M; j++)
        h[2 * (size_t)M + j] = (double)sinl(3.14159265358979323846264338327950288L * (long double)j / (long double)M);
    size_t n = h.size();
    void *d = nullptr;
    int rc = dev_alloc(c, &d, n * (f32 ? 4 : 8));
    if (rc) return rc;
    if (f32) {
        std::vector<float> hf(h.begin(), h.end());
        HIPCHK(hipMemcpy(d, hf.data(), n * 4, hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemcpy(d, h.data(), n * 8, hipMemcpyHostToDevice));
    }
    m[M] = d;
    *out = d;
    return HH_OK;
}

// DST variant for lines of length N.  Automatic choice: the half-length
// transform at every power-of-two N+1: 20-45% faster than the odd extension
// above N+1 = 1024 (503 against 636 us for 1 GB of rows at N+1 = 2048, 566
// against 1055 at 4096), and accurate to 7e-15 relative per transform at
// N+1 = 4096; its error is correlated across the spectrum (prefix sums of
// transform values), so a transform and its inverse give 7e-14 there
// against 5e-16 for the odd extension, far inside a solve's 1e-12.
static int dst_algo_for(const hh_context *c, int N) {
    if (c->dst_algo >= 0) return c->dst_algo;
    return 0;
}

// Block size of the automatic choice: one-wave blocks (one line each) for
// contiguous rows of N+1 >= 512, measured 3-8% faster there (N+1 = 512: 446
// against 484 us per GB of rows; 1024: 496 against 521) and slower at 256;
// 256-thread blocks for strided lines, whose tiles they widen; in FP32,
// 512-thread blocks (tiles of 32 lines at N+1 = 256: 69 against 74 us per
// 255^3 pass; in FP64 such a tile's LDS leaves one block per CU and it is
// slower, 120 against 116 us).
static int dst_algo_lines(const hh_context *c, const hh::Lines &ln, int tc) {
    int algo = dst_algo_for(c, ln.N);
    // one-wave blocks pay for FP64 rows only (N+1 = 1024: 472 against 521 us); with FP32 storage
    // 256-thread blocks stay faster (352 against 393 us per pass at 1023^2 x 64)
    if (c->dst_algo < 0 && !tc && ln.elem_stride == 1 && !ln.in_tw && !ln.out_tw && ln.N + 1 >= 512) algo |= 2;
    if (c->dst_algo < 0 && ln.elem_stride != 1 && tc && ln.N + 1 >= 128 && ln.N + 1 <= 256) algo |= 4;
    return algo;
}

// One DST pass over lines; in-place when the FFT path applies.
static int dst_pass(hh_context *c, const hh::Lines &ln0, int ti, int tc, int to, const void *in, void *out,
                    double acc = 0) {
    hh::Lines ln = ln0;
    ln.acc = acc;
    void *tw = nullptr, *S = nullptr;
    int rc;
    const int algo = dst_algo_lines(c, ln, tc);
    if (hh::fft_supported(ln.N))
        rc = (algo & 1) == 0 ? get_half_table(c, ln.N, tc, &tw) : get_twiddles(c, ln.N, tc, &tw);
    else rc = get_dst_matrix(c, ln.N, tc, &S);
    if (rc) return rc;
    const double elems = (double)ln.N * ln.ncols * ln.nplanes;
    TimeScope ts(c, tw ? 0 : 1, elems * ((ti ? 8 : 16) + (to ? 8 : 16) * (acc != 0 ? 2 : 1)));
    hh::DstTypes ty{ti, tc, to};
    ty.algo = algo;
    hipError_t e = hh::launch_dst(ln, ty, in, out, tw, S, c->stream);
    if (e != hipSuccess) return fail(HH_ERR_DEVICE, "DST launch (N=%d) failed: %s", ln.N, hipGetErrorString(e));
    return HH_OK;
}

static hh::Lines rows_lines(int N, long nrows) {
    hh::Lines l;
    l.N = N; l.ncols = nrows; l.nplanes = 1; l.plane_stride = (long)N * nrows; l.col_stride = N; l.elem_stride = 1;
    return l;
}

static hh::Lines cols_lines(int Nx, int Ny, long nplanes) {
    hh::Lines l;
    l.N = Ny; l.ncols = Nx; l.nplanes = nplanes; l.plane_stride = (long)Nx * Ny; l.col_stride = 1; l.elem_stride = Nx;
    return l;
}

static bool plan_all_fft(const hh_plan *p) {
    return hh::fft_supported(p->Nx) && (p->dim == 2 || hh::fft_supported(p->Ny));
}

// ------------------------------------------------------------------ errors
extern "C" const char *hh_last_error(void) { return g_err.c_str(); }

extern "C" const char *hh_status_string(int s) {
    switch (s) {
    case HH_OK: return "ok";
    case HH_ERR_ARG: return "invalid argument";
    case HH_ERR_DEVICE: return "device error";
    case HH_ERR_NOMEM: return "out of device memory";
    case HH_ERR_UNSUPPORTED: return "unsupported shape";
    case HH_ERR_SINGULAR: return "singular mode";
    default: return "unknown status";
    }
}

// ------------------------------------------------------------------ context
extern "C" int hh_context_create(int device, hh_context **out) {
    if (!out) return fail(HH_ERR_ARG, "hh_context_create: null output pointer");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(HH_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= n) return fail(HH_ERR_ARG, "device %d out of range (%d devices)", device, n);
    hh_context *c = new hh_context();
    c->device = device;
    HIPCHK(hipSetDevice(device));
    e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail(HH_ERR_DEVICE, "stream creation failed: %s", hipGetErrorString(e)); }
    c->stream = c->own;
    {
        // the static workspace of hh_solve_noalloc: its address on this device,
        // looked up once per device with the device current
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        auto it = g_noalloc.find(device);
        if (it == g_noalloc.end()) {
            void *ptr = nullptr;
            const hipError_t se = hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch));
            it = g_noalloc.emplace(device, std::make_pair(se == hipSuccess ? ptr : nullptr, se)).first;
        }
        c->noalloc_ws = it->second.first;
        c->noalloc_status = it->second.second;
    }
    if (hipMalloc((void **)&c->d_norm, hh::norm_buffer_doubles() * sizeof(double)) != hipSuccess) {
        (void)hipStreamDestroy(c->own);
        delete c;
        return fail(HH_ERR_NOMEM, "norm buffer allocation failed");
    }
    *out = c;
    return HH_OK;
}

extern "C" int hh_context_destroy(hh_context *c) {
    if (!c) return HH_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->tw64) (void)hipFree(kv.second);
    for (auto &kv : c->tw32) (void)hipFree(kv.second);
    for (auto &kv : c->S64) (void)hipFree(kv.second);
    for (auto &kv : c->S32) (void)hipFree(kv.second);
    for (auto &kv : c->th64) (void)hipFree(kv.second);
    for (auto &kv : c->th32) (void)hipFree(kv.second);
    if (c->scratch) (void)hipFree(c->scratch);
    for (auto &r : c->recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    (void)hipFree(c->d_norm);
    (void)hipStreamDestroy(c->own);
    delete c;
    return HH_OK;
}

extern "C" int hh_context_set_stream(hh_context *c, void *stream) {
    if (!c) return fail(HH_ERR_ARG, "null context");
    // NULL is the HIP default (null) stream, which orders against the work
    // of callers (e.g. PyTorch's default stream) that use it
    c->stream = (hipStream_t)stream;
    return HH_OK;
}

extern "C" void *hh_context_get_stream(hh_context *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int hh_context_synchronize(hh_context *c) {
    if (!c) return fail(HH_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return HH_OK;
}

extern "C" int hh_context_set_memory_budget(hh_context *c, size_t bytes) {
    if (!c) return fail(HH_ERR_ARG, "null context");
    c->budget = bytes;
    return HH_OK;
}

extern "C" size_t hh_context_bytes_in_use(const hh_context *c) { return c ? c->in_use : 0; }

static int collect_timing(hh_context *c) {
    if (c->recs.empty()) return HH_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto &r : c->recs) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        c->acc_ms[r.cat] += ms;
        c->acc_bytes[r.cat] += r.bytes;
        c->acc_n[r.cat] += 1;
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->recs.clear();
    return HH_OK;
}

extern "C" int hh_context_set_dst_algorithm(hh_context *c, int algo) {
    if (!c) return fail(HH_ERR_ARG, "null context");
    if (algo < -1 || algo > 15) return fail(HH_ERR_ARG, "DST algorithm must be -1..15 (got %d)", algo);
    std::lock_guard<std::mutex> lk(c->mu);
    c->dst_algo = algo;
    return HH_OK;
}

extern "C" int hh_context_enable_timing(hh_context *c, int on) {
    if (!c) return fail(HH_ERR_ARG, "null context");
    std::lock_guard<std::mutex> lk(c->mu);
    c->timing = on != 0;
    return HH_OK;
}

extern "C" int hh_context_read_timing(hh_context *c, double ms[5], long launches[5], double bytes[5]) {
    if (!c) return fail(HH_ERR_ARG, "null context");
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = collect_timing(c);
    if (rc) return rc;
    for (int k = 0; k < 5; k++) {
        if (ms) ms[k] = c->acc_ms[k];
        if (launches) launches[k] = c->acc_n[k];
        if (bytes) bytes[k] = c->acc_bytes[k];
    }
    return HH_OK;
}

extern "C" int hh_context_reset_timing(hh_context *c) {
    if (!c) return fail(HH_ERR_ARG, "null context");
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = collect_timing(c);
    for (int k = 0; k < 5; k++) { c->acc_ms[k] = 0; c->acc_bytes[k] = 0; c->acc_n[k] = 0; }
    return rc;
}

// ------------------------------------------------------------------ coefficients
static inline cd ld(const double *p, long i) { return p ? cd(p[2 * i], p[2 * i + 1]) : cd(0, 0); }
static inline void st(double *p, long i, cd v) { p[2 * i] = v.real(); p[2 * i + 1] = v.imag(); }

extern "C" int hh_coefficients2d(int order, int Nx, int Ny, double hx, double hy, const double *k2e,
                                 const double *k2y, const double *k2yy, double *bp, double *ap, double *b, double *a,
                                 double *bm, double *am) {
    if (Nx < 1 || Ny < 1 || !k2e || !bp || !ap || !b || !a || !bm || !am)
        return fail(HH_ERR_ARG, "hh_coefficients2d: bad arguments");
    if (order != 2 && order != 4 && order != 6) return fail(HH_ERR_ARG, "order must be 2, 4 or 6 (got %d)", order);
    const double hy2 = hy * hy, hy3 = hy2 * hy, hy4 = hy2 * hy2, Ryx = hy * hy / hx / hx;
    for (int j = 0; j < Ny; j++) {
        cd k2 = ld(k2e, j + 1), k2p = ld(k2e, j + 2), k2m = ld(k2e, j);
        if (order == 6) {                                   // coefficients.c:44-61
            cd ky2 = ld(k2y, j), kyy = ld(k2yy, j);
            st(ap, j, 2.0 / 3.0 + hy2 * k2p / 90.0 + hy3 * ky2 * (hy2 * k2p / 6.0 + 2.0 / 3.0) / 20.0);
            st(bp, j, 1.0 / 6.0 + hy2 * k2p / 90.0 + hy3 * ky2 / 120.0);
            st(a, j, -10.0 / 3.0 + 41.0 * hy2 * k2 / 45.0 - hy4 * k2 * k2 / 20.0 + hy4 * kyy / 20.0);
            st(b, j, 2.0 / 3.0 + hy2 * k2 / 90.0);
            st(am, j, 2.0 / 3.0 + hy2 * k2m / 90.0 - hy3 * ky2 * (hy2 * k2m / 6.0 + 2.0 / 3.0) / 20.0);
            st(bm, j, 1.0 / 6.0 + hy2 * k2m / 90.0 - hy3 * ky2 / 120.0);
        } else if (order == 4) {                            // coefficients.c:25-42
            st(bp, j, cd(Ryx / 12.0 + 1.0 / 12.0));
            st(ap, j, 5.0 / 6.0 - Ryx / 6.0 + hy2 * k2p / 12.0);
            st(b, j, Ryx * (5.0 / 6.0) - 1.0 / 6.0 + hy2 * k2 / 12.0);
            st(a, j, -Ryx * (5.0 / 3.0) - 5.0 / 3.0 + 2.0 * hy2 * k2 / 3.0);
            st(bm, j, cd(Ryx / 12.0 + 1.0 / 12.0));
            st(am, j, 5.0 / 6.0 - Ryx / 6.0 + hy2 * k2m / 12.0);
        } else {                                            // coefficients.c:9-23
            st(bp, j, cd(0));
            st(ap, j, cd(1));
            st(b, j, cd(Ryx));
            st(a, j, -2.0 * Ryx - 2.0 + k2 * hy2);
            st(bm, j, cd(0));
            st(am, j, cd(1));
        }
    }
    return HH_OK;
}

// compact (1, diag, 1) solve of a complex vector (4th_order_partials.c:322-336)
static void compact_line(std::vector<cd> &v, double diag) {
    int n = (int)v.size();
    std::vector<double> U(n);
    U[0] = diag;
    for (int j = 1; j < n; j++) U[j] = diag - 1.0 / U[j - 1];
    for (int j = 1; j < n; j++) v[j] = v[j] - (1.0 / U[j - 1]) * v[j - 1];
    v[n - 1] = v[n - 1] / U[n - 1];
    for (int j = n - 2; j >= 0; j--) v[j] = (v[j] - v[j + 1]) / U[j];
}

extern "C" int hh_coefficients3d(int order, int Nx, int Ny, int Nz, double hx, double hy, double hz,
                                 const double *k2e, const double *k2d, double *const coef[12]) {
    if (Nx < 1 || Ny < 1 || Nz < 1 || !k2e || !coef) return fail(HH_ERR_ARG, "hh_coefficients3d: bad arguments");
    for (int q = 0; q < 12; q++)
        if (!coef[q]) return fail(HH_ERR_ARG, "hh_coefficients3d: null coefficient array %d", q);
    if (order != 2 && order != 4 && order != 6) return fail(HH_ERR_ARG, "order must be 2, 4 or 6 (got %d)", order);
    double *A = coef[0], *B = coef[1], *C = coef[2], *D = coef[3], *Ap = coef[4], *Bp = coef[5], *Cp = coef[6],
           *Dp = coef[7], *Am = coef[8], *Bm = coef[9], *Cm = coef[10], *Dm = coef[11];
    const double Rzx = hz * hz / hx / hx, Rzy = hz * hz / hy / hy, h2 = hz * hz;
    std::vector<cd> k2(Nz + 2);
    for (int l = 0; l < Nz + 2; l++) k2[l] = ld(k2e, l);
    if (order == 6) {                                       // coefficients.c:67-103
        std::vector<cd> kp(Nz), kpp(Nz);
        for (int l = 0; l < Nz; l++) {
            kp[l] = 3.0 * (k2[l + 2] - k2[l]) / hz;
            kpp[l] = 12.0 * (k2[l] - 2.0 * k2[l + 1] + k2[l + 2]) / hz / hz;
        }
        kp[0] -= ld(k2d, 0); kp[Nz - 1] -= ld(k2d, 1);
        kpp[0] -= ld(k2d, 2); kpp[Nz - 1] -= ld(k2d, 3);
        compact_line(kp, 4.0);
        compact_line(kpp, 10.0);
        for (int i = 0; i < Nz; i++) {
            st(A, i, 1.0 / 10.0 + h2 * k2[i + 1] / 90.0);
            st(B, i, 7.0 / 15.0 - h2 * k2[i + 1] / 90.0);
            st(C, i, 7.0 / 15.0 - h2 * k2[i + 1] / 90.0);
            st(D, i, -64.0 / 15.0 + 14.0 * h2 * k2[i + 1] / 15.0 - h2 * h2 * k2[i + 1] * k2[i + 1] / 20.0 +
                         h2 * h2 * kpp[i] / 20.0);
            st(Ap, i, cd(1.0 / 30.0));
            st(Bp, i, 1.0 / 10.0 + h2 * k2[i + 2] / 90.0 + (kp[i] * h2 * hz / 120.0));
            st(Cp, i, 1.0 / 10.0 + h2 * k2[i + 2] / 90.0 + (kp[i] * h2 * hz / 120.0));
            st(Dp, i, 7.0 / 15.0 - h2 * k2[i + 2] / 90.0 + (kp[i] * h2 * hz / 20.0) * (1.0 / 3.0 + k2[i + 2] * h2 / 6.0));
            st(Am, i, cd(1.0 / 30.0));
            st(Bm, i, 1.0 / 10.0 + h2 * k2[i] / 90.0 - (kp[i] * h2 * hz / 120.0));
            st(Cm, i, 1.0 / 10.0 + h2 * k2[i] / 90.0 - (kp[i] * h2 * hz / 120.0));
            st(Dm, i, 7.0 / 15.0 - h2 * k2[i] / 90.0 - (kp[i] * h2 * hz / 20.0) * (1.0 / 3.0 + k2[i] * h2 / 6.0));
        }
    } else if (order == 4) {                                // coefficients.c:35-65
        for (int i = 0; i < Nz; i++) {
            st(A, i, cd((1.0 / 12.0) * (Rzx + Rzy)));
            st(B, i, (1.0 / 6.0) * (4.0 * Rzx - Rzy - 1.0 + 0.5 * h2 * k2[i + 1]));
            st(C, i, (1.0 / 6.0) * (4.0 * Rzy - Rzx - 1.0 + 0.5 * h2 * k2[i + 1]));
            st(D, i, -4.0 * (1.0 + Rzx + Rzy) / 3.0 + 0.5 * h2 * k2[i + 1]);
            st(Ap, i, cd(0));
            st(Bp, i, cd((1.0 / 12.0) * (1.0 + Rzx)));
            st(Cp, i, cd((1.0 / 12.0) * (1.0 + Rzy)));
            st(Dp, i, 1.0 - (1.0 / 6.0) * (2.0 + Rzx + Rzy) + (1.0 / 12.0) * h2 * k2[i + 2]);
            st(Am, i, cd(0));
            st(Bm, i, cd((1.0 / 12.0) * (1.0 + Rzx)));
            st(Cm, i, cd((1.0 / 12.0) * (1.0 + Rzy)));
            st(Dm, i, 1.0 - (1.0 / 6.0) * (2.0 + Rzx + Rzy) + (1.0 / 12.0) * h2 * k2[i]);
        }
    } else {                                                // coefficients.c:7-33
        for (int i = 0; i < Nz; i++) {
            st(A, i, cd(0)); st(B, i, cd(Rzx)); st(C, i, cd(Rzy));
            st(D, i, -2.0 * (Rzx + Rzy + 1.0) + h2 * k2[i + 1]);
            st(Ap, i, cd(0)); st(Bp, i, cd(0)); st(Cp, i, cd(0)); st(Dp, i, cd(1));
            st(Am, i, cd(0)); st(Bm, i, cd(0)); st(Cm, i, cd(0)); st(Dm, i, cd(1));
        }
    }
    return HH_OK;
}

// ------------------------------------------------------------------ Robin closure
// Boundary condition du/dn + alpha u = g on the first (lo) or last (hi) row of
// the tridiagonal axis, the row itself lying on the boundary.  The central
// difference (u_1 - u_{-1}) / 2h places a ghost row u_{-1} = u_1 - 2h alpha u_0
// + 2h g outside the grid; substituting it into the boundary row's stencil
// M u_{-1} + C u_0 + P u_1 = F_0 gives (C - 2h alpha M) u_0 + (P + M) u_1 =
// F_0 - 2h M g (mirrored on the hi side).  M, C and P are the same
// combinations of the DST-direction neighbour sums in every row, so the fold
// is band-wise on the coefficient vectors, and the DST still diagonalises
// the folded operator.  corr receives -2h M g (the row of F to add).
// bands: pointers to nb centre, nb plus and nb minus vectors, in that order.
static void robin_fold(int n, double h, cd alpha, bool lo, double *const *ctr, double *const *pls,
                       double *const *mns, int nb, cd *orig) {
    const int r = lo ? 0 : n - 1;
    for (int q = 0; q < nb; q++) {
        const cd m = ld(mns[q], r), p = ld(pls[q], r), c = ld(ctr[q], r);
        const cd out = lo ? m : p;             // band that reached the ghost row
        orig[q] = out;
        st(ctr[q], r, c - 2.0 * h * alpha * out);
        if (lo) { st(pls[q], r, p + m); st(mns[q], r, cd(0)); }
        else    { st(mns[q], r, m + p); st(pls[q], r, cd(0)); }
    }
}

extern "C" int hh_robin2d(int Nx, int Ny, double h, const double *alpha_lo, const double *g_lo,
                          const double *alpha_hi, const double *g_hi, double *const coef[6], double *corr_lo,
                          double *corr_hi) {
    if (Nx < 1 || Ny < 2 || !(h > 0) || !coef) return fail(HH_ERR_ARG, "hh_robin2d: bad arguments");
    for (int q = 0; q < 6; q++)
        if (!coef[q]) return fail(HH_ERR_ARG, "hh_robin2d: null coefficient array %d", q);
    // [bp, ap, b, a, bm, am]: centre (b, a), plus (bp, ap), minus (bm, am); per band the x-operator
    // is (first) * (u_{i-1} + u_{i+1}) + (second) * u_i
    double *ctr[2] = {coef[2], coef[3]}, *pls[2] = {coef[0], coef[1]}, *mns[2] = {coef[4], coef[5]};
    for (int side = 0; side < 2; side++) {
        const double *al = side == 0 ? alpha_lo : alpha_hi, *g = side == 0 ? g_lo : g_hi;
        double *corr = side == 0 ? corr_lo : corr_hi;
        if (!al) continue;
        cd o[2];
        robin_fold(Ny, h, cd(al[0], al[1]), side == 0, ctr, pls, mns, 2, o);
        if (!corr) continue;
        for (int i = 0; i < Nx; i++) {
            const cd nb = (i > 0 ? ld(g, i - 1) : cd(0)) + (i + 1 < Nx ? ld(g, i + 1) : cd(0));
            st(corr, i, -2.0 * h * (o[0] * nb + o[1] * ld(g, i)));
        }
    }
    return HH_OK;
}

extern "C" int hh_robin3d(int Nx, int Ny, int Nz, double h, const double *alpha_lo, const double *g_lo,
                          const double *alpha_hi, const double *g_hi, double *const coef[12], double *corr_lo,
                          double *corr_hi) {
    if (Nx < 1 || Ny < 1 || Nz < 2 || !(h > 0) || !coef) return fail(HH_ERR_ARG, "hh_robin3d: bad arguments");
    for (int q = 0; q < 12; q++)
        if (!coef[q]) return fail(HH_ERR_ARG, "hh_robin3d: null coefficient array %d", q);
    // [a, b, c, d | ap..dp | am..dm]: per band the (x, y)-operator is a * (diagonal neighbours)
    // + b * (x neighbours) + c * (y neighbours) + d * u
    double *ctr[4] = {coef[0], coef[1], coef[2], coef[3]}, *pls[4] = {coef[4], coef[5], coef[6], coef[7]},
           *mns[4] = {coef[8], coef[9], coef[10], coef[11]};
    auto G = [&](const double *g, int i, int j) {
        return (i >= 0 && i < Nx && j >= 0 && j < Ny) ? ld(g, i + (long)j * Nx) : cd(0);
    };
    for (int side = 0; side < 2; side++) {
        const double *al = side == 0 ? alpha_lo : alpha_hi, *g = side == 0 ? g_lo : g_hi;
        double *corr = side == 0 ? corr_lo : corr_hi;
        if (!al) continue;
        cd o[4];
        robin_fold(Nz, h, cd(al[0], al[1]), side == 0, ctr, pls, mns, 4, o);
        if (!corr) continue;
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                const cd dg = G(g, i - 1, j - 1) + G(g, i + 1, j - 1) + G(g, i - 1, j + 1) + G(g, i + 1, j + 1);
                const cd xg = G(g, i - 1, j) + G(g, i + 1, j), yg = G(g, i, j - 1) + G(g, i, j + 1);
                st(corr, i + (long)j * Nx, -2.0 * h * (o[0] * dg + o[1] * xg + o[2] * yg + o[3] * G(g, i, j)));
            }
    }
    return HH_OK;
}

// ------------------------------------------------------------------ plans
// Device band-table order from the caller's operator vectors.
// 2D: input [op][bp, ap, b, a, bm, am][Ny] -> table [op][q](bm, am, b, a, bp, ap);
// 3D: input [op][a..dm][Nz] -> table [op][l][12].
static void fill_real(hh_plan *p) {
    const size_t cnt = p->hcoef.size();
    p->hr64.resize(cnt);
    p->hr32.resize(cnt);
    int real = 1;
    for (size_t i = 0; i < cnt; i++) {
        p->hr64[i] = p->hcoef[i].real();
        p->hr32[i] = (float)p->hcoef[i].real();
        if (p->hcoef[i].imag() != 0.0) real = 0;
    }
    p->real_op = real;
}

static void build_table(int dim, int n, int nop, const double *coef, std::vector<cd> &tab) {
    if (dim == 2) {
        static const int src[6] = {4, 5, 2, 3, 0, 1};
        tab.resize(6 * (size_t)n * nop);
        for (int o = 0; o < nop; o++)
            for (int q = 0; q < n; q++)
                for (int t = 0; t < 6; t++)
                    tab[((size_t)o * n + q) * 6 + t] = ld(coef, ((size_t)o * 6 + src[t]) * n + q);
    } else {
        tab.resize(12 * (size_t)n * nop);
        for (int o = 0; o < nop; o++)
            for (int l = 0; l < n; l++)
                for (int q = 0; q < 12; q++)
                    tab[((size_t)o * n + l) * 12 + q] = ld(coef, ((size_t)o * 12 + q) * n + l);
    }
}

static void fill_table(hh_plan *p, int nop, const double *coef) {
    build_table(p->dim, p->dim == 2 ? p->Ny : p->Nz, nop, coef, p->hcoef);
    fill_real(p);
    hh::classify_modes(p->dim, p->Nx, p->Ny, p->Nz, nop, p->hcoef.data(), p->hlamx.data(), p->hlamy.data(),
                       p->real_op != 0, p->kappa_flag, p->ms);
}

// eigenvalues 2cos(k pi/(N+1)), k = 1..N (LU.c:12-14)
static void eigenvalues(int N, std::vector<double> &v) {
    v.resize(N);
    for (int k = 0; k < N; k++) v[k] = 2.0 * std::cos((k + 1) * M_PI / (N + 1));
}

// Device array holding at least `bytes`; a new allocation bumps the plan's
// generation (recorded graphs hold the old address).
static int ensure_dev(hh_plan *p, void **ptr, size_t *cap, size_t bytes) {
    if (*ptr && bytes <= *cap) return HH_OK;
    hh_context *c = p->ctx;
    const size_t old = *ptr ? *cap : 0;
    if (*ptr) {
        HIPCHK(hipStreamSynchronize(c->stream));
        dev_free(c, *ptr, *cap);
        *ptr = nullptr;
        *cap = 0;
    }
    // grow by at least half again (a serving loop's operators flag a few
    // modes more or fewer from batch to batch: reallocate seldom, since a
    // free waits for the stream)
    size_t b = bytes ? bytes : 16;
    if (old) b = std::max(b, old + old / 2);
    int rc = dev_alloc(c, ptr, b);
    if (rc) return rc;
    *cap = b;
    p->gen++;
    return HH_OK;
}

// Copies the near-resonant mode tables to the device, stream-ordered after
// the work already queued (the host vectors stay untouched until the next
// fill_table, which waits for the stream first).
static int upload_modes(hh_plan *p) {
    hh_context *c = p->ctx;
    hh::ModeSet &m = p->ms;
    int rc;
    // device layout of the factors: [row][entry](L, P, R, Up), so that the
    // lanes of neighbouring entries read neighbouring records
    const long ne = (long)m.fmode.size(), n = m.n;
    p->hfacc_t.resize(m.facc.size());
    p->hfacr_t.resize(m.facr.size());
    for (long e = 0; e < ne; e++)
        for (long q = 0; q < n; q++)
            for (int k = 0; k < 4; k++) {
                p->hfacc_t[(q * ne + e) * 4 + k] = m.facc[(e * n + q) * 4 + k];
                if (!m.facr.empty()) p->hfacr_t[(q * ne + e) * 4 + k] = m.facr[(e * n + q) * 4 + k];
            }
    if ((rc = ensure_dev(p, (void **)&p->d_foff, &p->cap_foff, m.foff.size() * sizeof(int)))) return rc;
    if ((rc = ensure_dev(p, (void **)&p->d_fmode, &p->cap_fmode, m.fmode.size() * sizeof(int)))) return rc;
    if ((rc = ensure_dev(p, (void **)&p->d_bits, &p->cap_bits, m.bits.size() * sizeof(uint32_t)))) return rc;
    if ((rc = ensure_dev(p, &p->d_facc, &p->cap_facc, m.facc.size() * sizeof(cd)))) return rc;
    if ((rc = ensure_dev(p, &p->d_facr, &p->cap_facr, m.facr.size() * sizeof(double)))) return rc;
    auto cp = [&](void *d, const void *h, size_t bytes) -> int {
        if (bytes) HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
        return HH_OK;
    };
    if ((rc = cp(p->d_foff, m.foff.data(), m.foff.size() * sizeof(int)))) return rc;
    if ((rc = cp(p->d_fmode, m.fmode.data(), m.fmode.size() * sizeof(int)))) return rc;
    if ((rc = cp(p->d_bits, m.bits.data(), m.bits.size() * sizeof(uint32_t)))) return rc;
    if ((rc = cp(p->d_facc, p->hfacc_t.data(), p->hfacc_t.size() * sizeof(cd)))) return rc;
    return cp(p->d_facr, p->hfacr_t.data(), p->hfacr_t.size() * sizeof(double));
}

static int upload_plan(hh_plan *p, int ncoef, int n) {
    hh_context *c = p->ctx;
    size_t b64 = (size_t)ncoef * n * p->nop * 16, b32 = (size_t)ncoef * n * p->nop * 8;
    int rc = dev_alloc(c, &p->coef64, b64);
    if (rc) return rc;
    rc = dev_alloc(c, &p->coef32, b32);
    if (rc) return rc;
    HIPCHK(hipMemcpy(p->coef64, p->hcoef.data(), b64, hipMemcpyHostToDevice));
    std::vector<float> h32(2 * (size_t)ncoef * n * p->nop);
    for (size_t i = 0; i < (size_t)ncoef * n * p->nop; i++) {
        h32[2 * i] = (float)p->hcoef[i].real();
        h32[2 * i + 1] = (float)p->hcoef[i].imag();
    }
    HIPCHK(hipMemcpy(p->coef32, h32.data(), b32, hipMemcpyHostToDevice));
    if ((rc = dev_alloc(c, &p->coefr64, b64 / 2))) return rc;
    if ((rc = dev_alloc(c, &p->coefr32, b32 / 2))) return rc;
    HIPCHK(hipMemcpy(p->coefr64, p->hr64.data(), b64 / 2, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p->coefr32, p->hr32.data(), b32 / 2, hipMemcpyHostToDevice));
    const std::vector<double> &lx = p->hlamx, &ly = p->hlamy;
    rc = dev_alloc(c, (void **)&p->lamx, lx.size() * 8);
    if (rc) return rc;
    HIPCHK(hipMemcpy(p->lamx, lx.data(), lx.size() * 8, hipMemcpyHostToDevice));
    if (p->dim == 3) {
        rc = dev_alloc(c, (void **)&p->lamy, ly.size() * 8);
        if (rc) return rc;
        HIPCHK(hipMemcpy(p->lamy, ly.data(), ly.size() * 8, hipMemcpyHostToDevice));
    }
    if ((rc = upload_modes(p))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return HH_OK;
}

// ------------------------------------------------------------------ host-only mode classification
static bool table_real(const std::vector<cd> &tab) {
    for (const cd &v : tab)
        if (v.imag() != 0.0) return false;
    return true;
}

extern "C" int hh_operator_modes(int dim, int Nx, int Ny, int Nz, const double *coef, double kappa, int *nflagged,
                                 int *modes, int cap, int *singular, double *kappa_worst) {
    if ((dim != 2 && dim != 3) || Nx < 2 || Ny < 2 || (dim == 3 && Nz < 2) || !coef || !(kappa >= 0) ||
        (cap > 0 && !modes))
        return fail(HH_ERR_ARG, "hh_operator_modes: bad arguments");
    std::vector<cd> tab;
    build_table(dim, dim == 2 ? Ny : Nz, 1, coef, tab);
    std::vector<double> lx, ly;
    eigenvalues(Nx, lx);
    if (dim == 3) eigenvalues(Ny, ly);
    hh::ModeSet ms;
    hh::classify_modes(dim, Nx, Ny, Nz, 1, tab.data(), lx.data(), ly.data(), table_real(tab), kappa, ms);
    if (nflagged) *nflagged = (int)ms.fmode.size();
    for (int i = 0; i < (int)ms.fmode.size() && i < cap; i++) modes[i] = ms.fmode[i];
    if (singular) *singular = ms.singular[0];
    if (kappa_worst) *kappa_worst = ms.kappa_worst[0];
    return HH_OK;
}

extern "C" int hh_operator_mode_factors(int dim, int Nx, int Ny, int Nz, const double *coef, int mode,
                                        double *factors) {
    const long nmodes = dim == 2 ? Nx : (long)Nx * Ny;
    if ((dim != 2 && dim != 3) || Nx < 2 || Ny < 2 || (dim == 3 && Nz < 2) || !coef || !factors || mode < 0 ||
        mode >= nmodes)
        return fail(HH_ERR_ARG, "hh_operator_mode_factors: bad arguments");
    const int n = dim == 2 ? Ny : Nz;
    std::vector<cd> tab;
    build_table(dim, n, 1, coef, tab);
    std::vector<double> lx, ly;
    eigenvalues(Nx, lx);
    if (dim == 3) eigenvalues(Ny, ly);
    // classify with kappa = 0: every mode flagged, entry index = mode index
    hh::ModeSet ms;
    const bool real = table_real(tab);
    hh::classify_modes(dim, Nx, Ny, Nz, 1, tab.data(), lx.data(), ly.data(), real, 0.0, ms);
    const cd *f = ms.facc.data() + (size_t)mode * n * 4;
    for (long q = 0; q < n; q++) {   // (L, R, Up) of each row
        const cd v[3] = {f[4 * q], f[4 * q + 2], f[4 * q + 3]};
        for (int k = 0; k < 3; k++) {
            factors[2 * (3 * q + k)] = v[k].real();
            factors[2 * (3 * q + k) + 1] = v[k].imag();
        }
    }
    return ms.singular[0] ? fail(HH_ERR_SINGULAR, "operator is singular to working precision") : HH_OK;
}

static int prepare_tables(hh_plan *p);

static int plan_check_dims(int dim, int Nx, int Ny, int Nz) {
    if (Nx < 2 || Ny < 2 || (dim == 3 && Nz < 2))
        return fail(HH_ERR_ARG, "grid sizes must be at least 2 (got %d x %d x %d)", Nx, Ny, Nz);
    long big = 1L << 30;
    if ((long)Nx * Ny * (dim == 3 ? Nz : 1) > 64 * big) return fail(HH_ERR_UNSUPPORTED, "grid too large");
    int sys = dim == 2 ? Ny : Nz;
    if (sys > 8192) return fail(HH_ERR_UNSUPPORTED, "system length %d above 8192", sys);
    return HH_OK;
}

extern "C" int hh_plan2d_create(hh_context *ctx, int Nx, int Ny, const double *bp, const double *ap,
                                const double *b, const double *a, const double *bm, const double *am, int precision,
                                hh_plan **out) {
    if (!ctx || !out || !bp || !ap || !b || !a || !bm || !am) return fail(HH_ERR_ARG, "hh_plan2d_create: null argument");
    if (Ny < 1) return fail(HH_ERR_ARG, "grid sizes must be at least 2 (got %d x %d)", Nx, Ny);
    const double *v[6] = {bp, ap, b, a, bm, am};
    std::vector<double> packed(2 * 6 * (size_t)Ny);
    for (int q = 0; q < 6; q++) memcpy(packed.data() + 2 * (size_t)q * Ny, v[q], 16 * (size_t)Ny);
    return hh_plan2d_create_multi(ctx, Nx, Ny, 1, packed.data(), precision, out);
}

extern "C" int hh_plan2d_create_multi(hh_context *ctx, int Nx, int Ny, int nop, const double *coef, int precision,
                                      hh_plan **out) {
    if (!ctx || !out || !coef) return fail(HH_ERR_ARG, "hh_plan2d_create_multi: null argument");
    if (nop < 1) return fail(HH_ERR_ARG, "number of operators must be >= 1 (got %d)", nop);
    if (precision != HH_PREC_F64 && precision != HH_PREC_MIXED) return fail(HH_ERR_ARG, "unknown precision %d", precision);
    int rc = plan_check_dims(2, Nx, Ny, 1);
    if (rc) return rc;
    if ((rc = set_device(ctx))) return rc;
    hh_plan *p = new hh_plan();
    p->ctx = ctx; p->dim = 2; p->Nx = Nx; p->Ny = Ny; p->Nz = 1; p->prec = precision; p->nop = nop;
    p->cap_nop = nop;
    eigenvalues(Nx, p->hlamx);
    fill_table(p, nop, coef);
    if ((rc = upload_plan(p, 6, Ny))) { hh_plan_destroy(p); return rc; }
    // the transform tables the plan's solves read: a solve then allocates nothing but workspace
    if ((rc = prepare_tables(p))) { hh_plan_destroy(p); return rc; }
    *out = p;
    return HH_OK;
}

extern "C" int hh_plan3d_create(hh_context *ctx, int Nx, int Ny, int Nz, const double *const coef[12], int precision,
                                hh_plan **out) {
    if (!ctx || !out || !coef) return fail(HH_ERR_ARG, "hh_plan3d_create: null argument");
    for (int q = 0; q < 12; q++)
        if (!coef[q]) return fail(HH_ERR_ARG, "hh_plan3d_create: null coefficient array %d", q);
    if (Nz < 1) return fail(HH_ERR_ARG, "grid sizes must be at least 2 (got Nz = %d)", Nz);
    std::vector<double> packed(2 * 12 * (size_t)Nz);
    for (int q = 0; q < 12; q++) memcpy(packed.data() + 2 * (size_t)q * Nz, coef[q], 16 * (size_t)Nz);
    return hh_plan3d_create_multi(ctx, Nx, Ny, Nz, 1, packed.data(), precision, out);
}

extern "C" int hh_plan3d_create_multi(hh_context *ctx, int Nx, int Ny, int Nz, int nop, const double *coef,
                                      int precision, hh_plan **out) {
    if (!ctx || !out || !coef) return fail(HH_ERR_ARG, "hh_plan3d_create_multi: null argument");
    if (nop < 1) return fail(HH_ERR_ARG, "number of operators must be >= 1 (got %d)", nop);
    if (precision != HH_PREC_F64 && precision != HH_PREC_MIXED) return fail(HH_ERR_ARG, "unknown precision %d", precision);
    int rc = plan_check_dims(3, Nx, Ny, Nz);
    if (rc) return rc;
    if ((rc = set_device(ctx))) return rc;
    hh_plan *p = new hh_plan();
    p->ctx = ctx; p->dim = 3; p->Nx = Nx; p->Ny = Ny; p->Nz = Nz; p->prec = precision; p->nop = nop;
    p->cap_nop = nop;
    eigenvalues(Nx, p->hlamx);
    eigenvalues(Ny, p->hlamy);
    fill_table(p, nop, coef);
    if ((rc = upload_plan(p, 12, Nz))) { hh_plan_destroy(p); return rc; }
    // the transform tables the plan's solves read: a solve then allocates nothing but workspace
    if ((rc = prepare_tables(p))) { hh_plan_destroy(p); return rc; }
    *out = p;
    return HH_OK;
}

extern "C" int hh_plan_num_operators(const hh_plan *p) { return p ? p->nop : 0; }

extern "C" int hh_plan_destroy(hh_plan *p) {
    if (!p) return HH_OK;
    hh_context *c = p->ctx;
    int nc = p->dim == 2 ? 6 : 12, n = p->dim == 2 ? p->Ny : p->Nz;
    (void)hipStreamSynchronize(c->stream);
    dev_free(c, p->coef64, (size_t)nc * n * p->cap_nop * 16);
    dev_free(c, p->coef32, (size_t)nc * n * p->cap_nop * 8);
    dev_free(c, p->coefr64, (size_t)nc * n * p->cap_nop * 8);
    dev_free(c, p->coefr32, (size_t)nc * n * p->cap_nop * 4);
    dev_free(c, p->lamx, (size_t)p->Nx * 8);
    if (p->lamy) dev_free(c, p->lamy, (size_t)p->Ny * 8);
    dev_free(c, p->d_foff, p->cap_foff);
    dev_free(c, p->d_fmode, p->cap_fmode);
    dev_free(c, p->d_bits, p->cap_bits);
    dev_free(c, p->d_facc, p->cap_facc);
    dev_free(c, p->d_facr, p->cap_facr);
    delete p;
    return HH_OK;
}

extern "C" hh_context *hh_plan_context(const hh_plan *p) { return p ? p->ctx : nullptr; }

extern "C" int hh_plan_dims(const hh_plan *p, int dims[3]) {
    if (!p || !dims) return fail(HH_ERR_ARG, "hh_plan_dims: null argument");
    dims[0] = p->Nx; dims[1] = p->Ny; dims[2] = p->Nz;
    return p->dim;
}

extern "C" long hh_plan_field_size(const hh_plan *p) {
    return p ? (long)p->Nx * p->Ny * (p->dim == 3 ? p->Nz : 1) : 0;
}

extern "C" size_t hh_plan_workspace_bytes(const hh_plan *p, int batch) {
    if (!p || batch < 1) return 0;
    size_t f = (size_t)hh_plan_field_size(p) * batch;
    size_t ws = 0;
    if (!plan_all_fft(p)) ws += f * 16;      // ping-pong buffer of the dense DST path
    else if (p->dim == 2 && p->tile_w != 0) ws += f * (p->prec == HH_PREC_MIXED ? 8 : 16);   // mode-tiled spectrum
    if (p->prec == HH_PREC_MIXED) ws += f * 8;   // FP32 correction
    return ws;
}

extern "C" int hh_plan_set_refine(hh_plan *p, double tol, int max_iter) {
    if (!p || tol < 0 || max_iter < 0) return fail(HH_ERR_ARG, "hh_plan_set_refine: bad arguments");
    p->gen++;
    p->tol = tol;
    p->max_iter = max_iter;
    return HH_OK;
}

extern "C" int hh_plan_set_tri_shape(hh_plan *p, int W, int rows_per_thread, int rhs_per_thread) {
    if (!p) return fail(HH_ERR_ARG, "null plan");
    p->gen++;
    if (W == 0 && rows_per_thread == 0) { p->tri_W = p->tri_mm = 0; p->tri_nr = 1; return HH_OK; }
    int n = p->dim == 2 ? p->Ny : p->Nz;
    if (W < 1 || W > 64 || (rows_per_thread != 4 && rows_per_thread != 8 && rows_per_thread != 16) ||
        (rhs_per_thread != 1 && rhs_per_thread != 2 && rhs_per_thread != 4))
        return fail(HH_ERR_ARG, "tridiagonal shape W=%d rows=%d rhs=%d not available", W, rows_per_thread,
                    rhs_per_thread);
    int P = (n + rows_per_thread - 1) / rows_per_thread;
    if (P > n / 2) P = n / 2;
    if (P < 1 || W * P > 1024 || (n + P - 1) / P > rows_per_thread)
        return fail(HH_ERR_UNSUPPORTED, "tridiagonal shape W=%d rows=%d does not fit n=%d", W, rows_per_thread, n);
    // the reduced system's two PCR buffers (A, C and the FP64 right-hand sides; A and C real for the
    // real-band kernel of real operators, complex otherwise) must fit the 160 KiB of LDS.  A shape that
    // fits only real bands is replaced by the built-in one while the operators are complex.
    if (hh::tri_lds_bytes(hh_plan_is_real_operator(p) ? 8 : 16, 16, P, W, rhs_per_thread) > (size_t)160 * 1024)
        return fail(HH_ERR_UNSUPPORTED, "tridiagonal shape W=%d rows=%d rhs=%d needs more LDS than a CU has", W,
                    rows_per_thread, rhs_per_thread);
    p->tri_W = W;
    p->tri_mm = rows_per_thread;
    p->tri_nr = rhs_per_thread;
    return HH_OK;
}

extern "C" int hh_plan_set_real_operator(hh_plan *p, int mode) {
    if (!p) return fail(HH_ERR_ARG, "null plan");
    p->gen++;
    if (mode < -1 || mode > 0) return fail(HH_ERR_ARG, "real-operator mode must be -1 (automatic) or 0 (complex bands)");
    p->real_mode = mode;
    return HH_OK;
}

extern "C" int hh_plan_is_real_operator(const hh_plan *p) { return p && p->real_op && p->real_mode != 0; }

extern "C" int hh_plan_set_tile_width(hh_plan *p, int tw) {
    if (!p) return fail(HH_ERR_ARG, "null plan");
    p->gen++;
    if (tw != -1 && (tw < 0 || tw > 64 || (tw & (tw - 1))))
        return fail(HH_ERR_ARG, "tile width must be -1, 0 or a power of two up to 64 (got %d)", tw);
    p->tile_w = tw;
    return HH_OK;
}

// Built-in shape of FP64 elimination on FP32 storage (mixed precision), on
// top of the FP64 one: in 3D from n = 255, 16 rows per thread (255^3: W = 16,
// 114 against 137 us per launch for 8 rows).
static void fp32_storage_rows(const hh_plan *p, int n, long nmodes, hh::TriHelm &t) {
    if (p->dim == 3 && t.mmax == 8 && n >= 255) {
        t.mmax = 16;
        hh::tri_shape_rows(n, nmodes, 16, &t.W, &t.P);
    }
}

// Shape of `rows` rows per thread and about `threads` threads per block:
// W the largest power of two with W P <= threads, in [wmin, 16], and no
// more than the modes need.
static void shape_for(int n, long nmodes, int rows, int threads, int wmin, hh::TriHelm &t) {
    int P = (n + rows - 1) / rows;
    if (P > n / 2) P = n / 2;
    if (P < 1) P = 1;
    int w = 1;
    while (w * 2 * P <= threads && w * 2 <= 16) w *= 2;
    if (w < wmin) w = wmin;
    while (w > 1 && w / 2 >= nmodes) w /= 2;
    t.W = w; t.P = P; t.mmax = rows;
}

// Built-in block shape of the mode solves (W, P, rows and right-hand sides
// per thread).  kind 0: FP64 elimination on FP64 storage, 1: FP64
// elimination on FP32 storage, 2: FP32 elimination; per_op: right-hand
// sides per operator in the batch.  Measured on MI355X (order-6 operators
// with k^2 = 2500 or 2500 + 50i, row-major spectrum; microseconds of the
// mode pass of one solve, near-resonant modes included; in brackets the
// former rule of 256-thread blocks, 16 rows per thread above n = 512):
//   real bands, FP64
//     2D  n <= 512: 16 rows, 2 rhs, ~512 threads: 127 x 256: 53.9 (60.0),
//         255 x 64: 59.2 (66.3), 511 x 64: 184 (216);
//         n <= 1024: 8 rows, 1 rhs, 512 threads: 1023 x 64: 632 (693)
//         (two blocks of 112 VGPRs per CU against one of 252);
//         above: 16 rows, 2 rhs, 512 threads: 2047 x 16: 848 (983)
//     3D  n <= 128: 8 rows, 2 rhs, 128 threads: 127^3 x 8: 138 (147);
//         above: 16 rows, 256 threads: 255^3 x 1: 187 (212)
//   complex bands, FP64
//         n <= 512: 8 rows, 2 rhs, 256 threads: 2D 127: 67 (75), 255: 79
//         (86), 511: 240 (265); 3D 127^3 x 8: 169 (228);
//         above: 16 rows, 1 rhs, W >= 4: 2047 x 16: 1292 (1321)
//   real bands, 2D, FP32 elimination: 16 rows, 512 threads: 1023 x 64:
//         449 (506), 2047 x 16: 609 (733), 511: 131 (133), 255: 51 (53)
// and the former rule elsewhere (within 2% of the best shape measured).
// Two right-hand sides per thread need two per operator in the batch.
static void builtin_tri_shape(const hh_plan *p, int kind, long per_op, bool real, hh::TriHelm &t) {
    const int n = p->dim == 2 ? p->Ny : p->Nz;
    const long nmodes = p->dim == 2 ? p->Nx : (long)p->Nx * p->Ny;
    hh::tri_shape(n, nmodes, kind == 0 ? hh::F64 : hh::F32, &t.W, &t.P, &t.mmax);
    t.nr = 1;
    if (kind == 1) {
        hh::tri_shape(n, nmodes, hh::F64, &t.W, &t.P, &t.mmax);
        fp32_storage_rows(p, n, nmodes, t);
        // real bands, 16 rows: two right-hand sides per thread share the
        // operator's pivots (N=1023 x 64: 419 against 531 us per launch)
        // when the grid still has 8 blocks per CU or more
        if (real && t.mmax == 16 && per_op >= 2) {
            const long tiles = (nmodes + t.W - 1) / t.W;
            if (tiles * p->nop * ((per_op + 1) / 2) >= 2048) t.nr = 2;
        }
        return;
    }
    if (n <= 16) return;
    int nr = 1;
    if (kind == 0 && real && p->dim == 2) {
        if (n <= 512) { shape_for(n, nmodes, 16, 512, 1, t); nr = 2; }
        else if (n <= 1024) shape_for(n, nmodes, 8, 512, 1, t);
        else { shape_for(n, nmodes, 16, 512, 1, t); nr = 2; }
    } else if (kind == 0 && real) {
        if (n <= 128) { shape_for(n, nmodes, 8, 128, 1, t); nr = 2; }
        else { shape_for(n, nmodes, 16, 256, 1, t); nr = 2; }
    } else if (kind == 0) {
        if (n <= 512) { shape_for(n, nmodes, 8, 256, 1, t); nr = 2; }
        else shape_for(n, nmodes, 16, 256, 4, t);
    } else if (real && p->dim == 2) {
        shape_for(n, nmodes, 16, 512, 1, t);
    }
    const size_t op = real ? 8 : 16;
    if (nr == 2 && per_op >= 2 && hh::tri_lds_bytes(op, 16, t.P, t.W, 2) <= 160 * 1024) t.nr = 2;
}

extern "C" int hh_plan_get_tri_shape(const hh_plan *p, int *W, int *P, int *rows_per_thread, int *rhs_per_thread) {
    if (!p) return fail(HH_ERR_ARG, "null plan");
    hh::TriHelm t;
    int n = p->dim == 2 ? p->Ny : p->Nz;
    // the shape of the solve's elimination (mixed precision: FP32), for a
    // batch of two or more right-hand sides per operator
    builtin_tri_shape(p, p->prec == HH_PREC_MIXED ? 2 : 0, 2, p->real_op && p->real_mode != 0, t);
    if (p->tri_W > 0) {   // the tuned shape, where its reduced system fits the LDS (as in solve_core)
        const bool mixed = p->prec == HH_PREC_MIXED, real = p->real_op && p->real_mode != 0;
        const int nr = mixed ? 1 : p->tri_nr;
        int P2 = (n + p->tri_mm - 1) / p->tri_mm;
        if (P2 > n / 2) P2 = n / 2;
        if (hh::tri_lds_bytes((real ? 8 : 16) / (mixed ? 2 : 1), mixed ? 8 : 16, P2, p->tri_W, nr) <=
            (size_t)160 * 1024) {
            t.W = p->tri_W;
            t.mmax = p->tri_mm;
            t.nr = nr;
            t.P = P2;
        }
    }
    if (W) *W = t.W;
    if (P) *P = t.P;
    if (rows_per_thread) *rows_per_thread = t.mmax;
    if (rhs_per_thread) *rhs_per_thread = t.nr;
    return HH_OK;
}

extern "C" int hh_plan_set_resonance_threshold(hh_plan *p, double kappa) {
    if (!p || !(kappa >= 0)) return fail(HH_ERR_ARG, "hh_plan_set_resonance_threshold: bad arguments");
    hh_context *c = p->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = set_device(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    const int old_nfx = p->ms.nfx;
    p->kappa_flag = kappa;
    hh::classify_modes(p->dim, p->Nx, p->Ny, p->Nz, p->nop, p->hcoef.data(), p->hlamx.data(), p->hlamy.data(),
                       p->real_op != 0, p->kappa_flag, p->ms);
    if (p->ms.nfx != old_nfx) p->gen++;
    return upload_modes(p);
}

extern "C" int hh_plan_mode_info(const hh_plan *p, int op, int *flagged, int *singular, double *kappa) {
    if (!p || op < 0 || op >= p->nop) return fail(HH_ERR_ARG, "hh_plan_mode_info: bad arguments");
    if (flagged) *flagged = p->ms.foff[op + 1] - p->ms.foff[op];
    if (singular) *singular = p->ms.singular[op];
    if (kappa) *kappa = p->ms.kappa_worst[op];
    return HH_OK;
}

extern "C" int hh_plan_flagged_modes(const hh_plan *p, int op, int *modes, int cap) {
    if (!p || op < 0 || op >= p->nop || (cap > 0 && !modes)) return fail(HH_ERR_ARG, "hh_plan_flagged_modes: bad arguments");
    const int n = p->ms.foff[op + 1] - p->ms.foff[op];
    for (int i = 0; i < n && i < cap; i++) modes[i] = p->ms.fmode[p->ms.foff[op] + i];
    return n;
}

extern "C" int hh_plan_last_refine(const hh_plan *p, int *sweeps, double *rel) {
    if (!p) return fail(HH_ERR_ARG, "null plan");
    if (sweeps) *sweeps = p->last_sweeps;
    if (rel) *rel = p->last_rel;
    return HH_OK;
}

// ------------------------------------------------------------------ solve
// Transform-solve-transform on one buffer.  Types: ti input, tc compute /
// intermediate storage, to output (0 = FP64, 1 = FP32).  tri_comp is the
// arithmetic of the tridiagonal elimination.
// acc != 0: sol += acc A^{-1} rhs, the last pass accumulating into the FP64
// sol (rhs is then consumed: it may serve as scratch).
static int solve_core(hh_plan *p, int ti, int tc, int to, int tri_comp, const void *rhs, void *sol, void *ws,
                      int batch, double acc = 0) {
    hh_context *c = p->ctx;
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    int rc;
    hh::TriHelm t;
    t.dim = p->dim; t.Nx = Nx; t.Ny = Ny; t.Nz = Nz;
    t.field = hh_plan_field_size(p);
    t.batch = batch;
    t.coef = tri_comp == 0 ? p->coef64 : p->coef32;
    t.real = p->real_op && p->real_mode != 0;
    t.coefr = tri_comp == 0 ? p->coefr64 : p->coefr32;
    t.nop = p->nop;
    t.cstride = p->dim == 2 ? 6L * Ny : 12L * Nz;
    t.lamx = p->lamx; t.lamy = p->lamy;
    int n = p->dim == 2 ? Ny : Nz;
    const long per_op = (batch + p->nop - 1) / p->nop;
    builtin_tri_shape(p, tri_comp == 1 ? 2 : tc == 1 ? 1 : 0, per_op, t.real, t);
    if (p->tri_W > 0 && p->tri_mm > 0) {   // shape chosen by the autotuner
        const int nr = tri_comp == 0 ? p->tri_nr : 1;
        int P = (n + p->tri_mm - 1) / p->tri_mm;
        if (P > n / 2) P = n / 2;
        const size_t op = (t.real ? 8 : 16) / (tri_comp == 0 ? 1 : 2), cx = tri_comp == 0 ? 16 : 8;
        if (hh::tri_lds_bytes(op, cx, P, p->tri_W, nr) <= (size_t)160 * 1024) {
            t.W = p->tri_W;
            t.mmax = p->tri_mm;
            t.nr = nr;
            t.P = P;
        }
    }
    // near-resonant modes: reference-order substitution from the host's
    // factors (FP64 arithmetic), then the partitioned kernel for the rest
    t.nfx = p->ms.nfx;
    if (t.nfx > 0) {
        t.fbits = p->d_bits;
        t.fwords = p->ms.words;
        t.foff = p->d_foff;
        t.fmode = p->d_fmode;
        t.fac = t.real ? p->d_facr : p->d_facc;
        t.nent = (long)p->ms.fmode.size();
    }
    // The two kernels touch disjoint outputs (the partitioned kernel reads the
    // near-resonant modes' rows but does not store them).  Run concurrently on
    // a second stream they took as long together (C2: 651 us) as one after
    // the other: the partitioned kernel's blocks fill the chip.
    auto tri = [&](void *buf) -> int {
        TimeScope ts(c, 2, (double)t.field * batch * (tc ? 16 : 32));
        hipError_t e = hh::launch_tri_seq(t, hh::DstTypes{tc, 0, tc}, buf, buf, c->stream);
        if (e != hipSuccess) return fail(HH_ERR_DEVICE, "near-resonant mode launch failed: %s", hipGetErrorString(e));
        e = hh::launch_tri_helm(t, hh::DstTypes{tc, tri_comp, tc}, buf, buf, c->stream);
        if (e != hipSuccess) return fail(HH_ERR_DEVICE, "tridiagonal launch failed: %s", hipGetErrorString(e));
        return HH_OK;
    };
    const bool fft = plan_all_fft(p);
    if (p->dim == 2) {
        // FFT path: the forward row DST writes the spectrum mode-tiled into
        // ws (tiles of W modes, the tridiagonal block width), the mode solves
        // run there in place, the inverse row DST reads it back row-major
        // into sol.  Dense path: ws is a plain row-major intermediate.
        hh::Lines lx = rows_lines(Nx, (long)Ny * batch);
        const int tw = p->tile_w < 0 ? t.W : p->tile_w;
        if (fft && tw > 0) {
            if (!ws || ws == rhs || ws == sol) return fail(HH_ERR_ARG, "internal: missing workspace");
            // tiles hold whole blocks: the built-in W (a power of two) is cut
            // to a narrower tile; a shape set by hand must fit as given
            if (p->tri_W == 0 && t.W > tw) t.W = tw;
            t.tiled = tw;
            hh::Lines fwd = lx, inv = lx;
            fwd.out_tw = tw; fwd.ny = Ny; fwd.field = t.field;
            inv.in_tw = tw; inv.ny = Ny; inv.field = t.field;
            if ((rc = dst_pass(c, fwd, ti, tc, tc, rhs, ws))) return rc;
            if ((rc = tri(ws))) return rc;
            return dst_pass(c, inv, tc, tc, to, ws, sol, acc);
        }
        // row-major spectrum: in sol (in place) on the FFT path, in ws otherwise
        void *mid = fft ? (tc == to && acc == 0 ? sol : ws) : ws;
        if (!mid) return fail(HH_ERR_ARG, "internal: missing workspace");
        if ((rc = dst_pass(c, lx, ti, tc, tc, rhs, mid))) return rc;
        if ((rc = tri(mid))) return rc;
        return dst_pass(c, lx, tc, tc, to, mid, sol, acc);
    }
    if (acc != 0 && (to != 0 || ti != tc)) return fail(HH_ERR_UNSUPPORTED, "internal: accumulating solve types");
    hh::Lines lx = rows_lines(Nx, (long)Ny * Nz * batch);
    hh::Lines ly = cols_lines(Nx, Ny, (long)Nz * batch);
    if (fft && tc == to && acc == 0) {
        if ((rc = dst_pass(c, lx, ti, tc, tc, rhs, sol))) return rc;
        if ((rc = dst_pass(c, ly, tc, tc, tc, sol, sol))) return rc;
        if ((rc = tri(sol))) return rc;
        if ((rc = dst_pass(c, ly, tc, tc, tc, sol, sol))) return rc;
        return dst_pass(c, lx, tc, tc, to, sol, sol);
    }
    if (!ws) return fail(HH_ERR_ARG, "internal: missing workspace");
    if (fft) {   // intermediate precision differs from the output: keep it in ws
        if ((rc = dst_pass(c, lx, ti, tc, tc, rhs, ws))) return rc;
        if ((rc = dst_pass(c, ly, tc, tc, tc, ws, ws))) return rc;
        if ((rc = tri(ws))) return rc;
        if ((rc = dst_pass(c, ly, tc, tc, tc, ws, ws))) return rc;
        return dst_pass(c, lx, tc, tc, to, ws, sol, acc);
    }
    // dense 3D path: ping-pong between ws and a second buffer holding the
    // intermediate in the compute type: sol (no smaller than the compute
    // type), or, when sol is accumulated into, the consumed rhs
    if (tc == 0 && to == 1) return fail(HH_ERR_UNSUPPORTED, "internal: FP64 intermediate in an FP32 output");
    void *mid = acc != 0 ? const_cast<void *>(rhs) : sol;
    if (mid == ws) return fail(HH_ERR_ARG, "internal: workspace aliases the intermediate");
    if ((rc = dst_pass(c, lx, ti, tc, tc, rhs, ws))) return rc;
    if ((rc = dst_pass(c, ly, tc, tc, tc, ws, mid))) return rc;
    if ((rc = tri(mid))) return rc;
    if ((rc = dst_pass(c, ly, tc, tc, tc, mid, ws))) return rc;
    return dst_pass(c, lx, tc, tc, to, ws, sol, acc);
}

static int residual_into(hh_plan *p, const void *u, const void *F, void *res, int res_f32, int batch) {
    TimeScope ts(p->ctx, 3, (double)hh_plan_field_size(p) * batch * (32 + (res_f32 ? 8 : 16)));
    hipError_t e;
    if (p->dim == 2)
        e = hh::launch_residual2d(p->Nx, p->Ny, batch, p->coef64, p->nop, u, F, res, res_f32, p->ctx->stream);
    else
        e = hh::launch_residual3d(p->Nx, p->Ny, p->Nz, batch, p->coef64, p->nop, u, F, res, res_f32, p->ctx->stream);
    if (e != hipSuccess) return fail(HH_ERR_DEVICE, "residual launch failed: %s", hipGetErrorString(e));
    return HH_OK;
}

static int norms_sync(hh_context *c, long n, const void *x, int x_f32, const void *y, double out[2], bool l2 = false) {
    hipError_t e = hh::launch_norms(n, x, x_f32, y, c->d_norm, c->stream, l2);
    if (e != hipSuccess) return fail(HH_ERR_DEVICE, "norm launch failed: %s", hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync(out, c->d_norm, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return HH_OK;
}

// HH_ERR_SINGULAR when an operator that right-hand sides 0..batch-1 use is
// singular to working precision.
static int check_singular(const hh_plan *p, int batch) {
    const int used = batch < p->nop ? batch : p->nop;
    for (int o = 0; o < used; o++)
        if (p->ms.singular[o])
            return fail(HH_ERR_SINGULAR, "operator %d is singular to working precision (a mode's condition "
                                         "estimate %.3g exceeds %.0e): k^2 is at or next to an eigenvalue",
                        o, p->ms.kappa_worst[o], hh::HH_SINGULAR_KAPPA);
    return HH_OK;
}

// hh_solve with the context lock held.
// A solve on workspace `scratch` (hh_plan_workspace_bytes(p, batch) bytes:
// the context's scratch, a graph's own workspace or the caller's).
static int solve_ws(hh_plan *p, const void *rhs, void *sol, int batch, void *scratch) {
    hh_context *c = p->ctx;
    int rc;
    const long n = hh_plan_field_size(p) * batch;
    if (p->prec == HH_PREC_F64) {
        void *ws = (plan_all_fft(p) && (p->dim == 3 || p->tile_w == 0)) ? nullptr : scratch;
        return solve_core(p, 0, 0, 0, 0, rhs, sol, ws, batch);
    }
    // mixed: FP32 transforms, storage and elimination of the well-separated
    // modes, FP64 substitution of the near-resonant ones, FP64 residual and
    // update; the correction lives in FP32 scratch, followed (2D) by the
    // FP32 mode-tiled spectrum of the solves
    // (or, with a dense DST, the FP32 intermediate of its out-of-place passes)
    void *dlt = scratch;
    const bool own_ws = !plan_all_fft(p) || (p->dim == 2 && p->tile_w != 0);
    void *tws = own_ws ? (void *)((char *)scratch + (size_t)n * 8) : dlt;
    if (rhs == sol) return fail(HH_ERR_ARG, "mixed precision solve needs d_sol != d_rhs");
    if ((rc = solve_core(p, 0, 1, 0, 1, rhs, sol, tws, batch))) return rc;   // u0 = A32^{-1} F
    if (p->tol == 0) {
        // fixed number of sweeps: no norms, no host synchronisation, so the
        // whole solve can be recorded into a HIP graph
        // (each sweep: r = A u - F, then u -= A32^{-1} r with the update
        // applied by the last DST pass as it stores)
        for (int it = 0; it < p->max_iter; it++) {
            if ((rc = residual_into(p, sol, rhs, dlt, 1, batch))) return rc;
            if ((rc = solve_core(p, 1, 1, 0, 1, dlt, sol, tws, batch, -1.0))) return rc;
        }
        p->last_sweeps = p->max_iter;
        p->last_rel = -1.0;
        return HH_OK;
    }
    double nf[2], nr[2];
    if ((rc = norms_sync(c, n, rhs, 0, nullptr, nf))) return rc;
    const double fn = nf[0] > 0 ? nf[0] : 1.0;
    int sweeps = 0;
    double rel = 0;
    for (;;) {
        // r = A u - F with FP64 arithmetic, stored in FP32 (its norm only
        // needs a few digits; the correction it drives is an FP32 solve)
        if ((rc = residual_into(p, sol, rhs, dlt, 1, batch))) return rc;
        if ((rc = norms_sync(c, n, dlt, 1, nullptr, nr))) return rc;
        rel = nr[0] / fn;
        if (rel <= p->tol || sweeps >= p->max_iter) break;
        if ((rc = solve_core(p, 1, 1, 0, 1, dlt, sol, tws, batch, -1.0))) return rc;   // u -= A32^{-1} r
        sweeps++;
    }
    p->last_sweeps = sweeps;
    p->last_rel = rel;
    return HH_OK;
}

static int solve_locked(hh_plan *p, const void *rhs, void *sol, int batch) {
    hh_context *c = p->ctx;
    int rc = check_singular(p, batch);
    if (rc) return rc;
    if ((rc = set_device(c))) return rc;
    size_t ws_bytes = hh_plan_workspace_bytes(p, batch);
    if (ws_bytes && (rc = ensure_scratch(c, ws_bytes))) return rc;
    return solve_ws(p, rhs, sol, batch, c->scratch);
}

extern "C" int hh_solve_noalloc(hh_plan *p, const void *rhs, void *sol, int batch) {
    if (!p || !rhs || !sol) return fail(HH_ERR_ARG, "hh_solve_noalloc: null argument");
    if (batch < 1) return fail(HH_ERR_ARG, "hh_solve_noalloc: batch must be >= 1 (got %d)", batch);
    hh_context *c = p->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_singular(p, batch);
    if (rc) return rc;
    if ((rc = set_device(c))) return rc;
    if (!c->noalloc_ws)
        return fail(HH_ERR_DEVICE, "hh_solve_noalloc: no address for the static workspace on device %d: %s", c->device,
                    hipGetErrorString(c->noalloc_status));
    const size_t ws = hh_plan_workspace_bytes(p, batch);
    if (ws > HH_NOALLOC_BYTES)
        return fail(HH_ERR_UNSUPPORTED, "hh_solve_noalloc: the solve needs %zu bytes of workspace, above the %zu of "
                                        "its static workspace (use hh_solve, or a smaller batch)",
                    ws, HH_NOALLOC_BYTES);
    c->no_alloc = true;
    Range range("hh_solve_noalloc");
    rc = solve_ws(p, rhs, sol, batch, c->noalloc_ws);
    c->no_alloc = false;
    return rc;
}

extern "C" int hh_solve(hh_plan *p, const void *rhs, void *sol, int batch) {
    if (!p || !rhs || !sol) return fail(HH_ERR_ARG, "hh_solve: null argument");
    if (batch < 1) return fail(HH_ERR_ARG, "hh_solve: batch must be >= 1 (got %d)", batch);
    std::lock_guard<std::mutex> lk(p->ctx->mu);
    Range range("hh_solve");
    return solve_locked(p, rhs, sol, batch);
}

// ------------------------------------------------------------------ graphs
struct hh_graph {
    hh_context *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int nodes = 0;
    // the graph's own workspace: later solves on the context may grow (and
    // move) the context's scratch, never this
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // plans and their generations at capture: a plan changed since (tables
    // reallocated, band kind, operator count or shapes) refuses the replay
    std::vector<std::pair<hh_plan *, long>> plans;
};

// Builds every table a plan's solves read (twiddles or DST matrices) so
// that recording a solve needs no allocation or synchronous copy.
static int prepare_tables(hh_plan *p) {
    hh_context *c = p->ctx;
    const int tc = p->prec == HH_PREC_F64 ? 0 : 1;
    const int lens[2] = {p->Nx, p->dim == 3 ? p->Ny : 0};
    for (int N : lens) {
        if (N <= 0) continue;
        void *t = nullptr;
        int rc;
        if (hh::fft_supported(N))
            rc = (dst_algo_for(c, N) & 1) == 0 ? get_half_table(c, N, tc, &t) : get_twiddles(c, N, tc, &t);
        else rc = get_dst_matrix(c, N, tc, &t);
        if (rc) return rc;
    }
    return HH_OK;
}

extern "C" int hh_graph_create(hh_context *c, int nsolves, hh_plan *const *plans, void *const *d_rhs,
                               void *const *d_sol, const int *batch, hh_graph **out) {
    if (!c || !out || nsolves < 1 || !plans || !d_rhs || !d_sol || !batch)
        return fail(HH_ERR_ARG, "hh_graph_create: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = set_device(c);
    if (rc) return rc;
    size_t ws = 0;
    for (int i = 0; i < nsolves; i++) {
        hh_plan *p = plans[i];
        if (!p || !d_rhs[i] || !d_sol[i] || batch[i] < 1) return fail(HH_ERR_ARG, "hh_graph_create: bad solve %d", i);
        if (p->ctx != c) return fail(HH_ERR_ARG, "hh_graph_create: plan %d belongs to another context", i);
        if (p->prec == HH_PREC_MIXED && p->tol != 0)
            return fail(HH_ERR_UNSUPPORTED, "hh_graph_create: mixed-precision plan %d needs fixed sweeps "
                                            "(hh_plan_set_refine(plan, 0, sweeps))", i);
        if ((rc = check_singular(p, batch[i]))) return rc;
        if ((rc = prepare_tables(p))) return rc;
        size_t b = hh_plan_workspace_bytes(p, batch[i]);
        if (b > ws) ws = b;
    }
    void *gws = nullptr;
    if (ws && (rc = dev_alloc(c, &gws, ws))) return rc;
    auto drop_ws = [&]() { if (gws) dev_free(c, gws, ws); };
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        drop_ws();
        return fail(HH_ERR_DEVICE, "hh_graph_create: stream synchronisation failed");
    }
    hipStream_t cap = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipStream_t saved = c->stream;
    const bool timing = c->timing;
    c->stream = cap;
    c->timing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int i = 0; i < nsolves && rc == HH_OK; i++) rc = solve_ws(plans[i], d_rhs[i], d_sol[i], batch[i], gws);
        hipError_t e2 = hipStreamEndCapture(cap, &g);
        if (rc == HH_OK && e2 != hipSuccess) rc = fail(HH_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e2));
    } else {
        rc = fail(HH_ERR_DEVICE, "hipStreamBeginCapture failed: %s", hipGetErrorString(e));
    }
    c->stream = saved;
    c->timing = timing;
    (void)hipStreamDestroy(cap);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        drop_ws();
        return rc;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        drop_ws();
        return fail(HH_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    hh_graph *G = new hh_graph();
    G->ctx = c;
    G->graph = g;
    G->exec = x;
    G->ws = gws;
    G->ws_bytes = ws;
    for (int i = 0; i < nsolves; i++) G->plans.emplace_back(plans[i], plans[i]->gen);
    size_t nn = 0;
    if (hipGraphGetNodes(g, nullptr, &nn) == hipSuccess) G->nodes = (int)nn;
    *out = G;
    return HH_OK;
}

extern "C" int hh_graph_launch(hh_graph *G) {
    if (!G) return fail(HH_ERR_ARG, "hh_graph_launch: null graph");
    hh_context *c = G->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = set_device(c);
    if (rc) return rc;
    for (const auto &pg : G->plans)
        if (pg.first->gen != pg.second)
            return fail(HH_ERR_ARG, "hh_graph_launch: a plan of the graph changed since it was recorded "
                                    "(tables reallocated, band kind, operator count or shapes); record it again");
    Range range("hh_graph_launch");
    HIPCHK(hipGraphLaunch(G->exec, c->stream));
    return HH_OK;
}

extern "C" int hh_graph_num_nodes(const hh_graph *G) { return G ? G->nodes : 0; }

extern "C" int hh_graph_destroy(hh_graph *G) {
    if (!G) return HH_OK;
    (void)hipStreamSynchronize(G->ctx->stream);
    if (G->exec) (void)hipGraphExecDestroy(G->exec);
    if (G->graph) (void)hipGraphDestroy(G->graph);
    if (G->ws) dev_free(G->ctx, G->ws, G->ws_bytes);
    delete G;
    return HH_OK;
}

extern "C" int hh_plan_set_operators(hh_plan *p, int nop, const double *coef) {
    if (!p || !coef) return fail(HH_ERR_ARG, "hh_plan_set_operators: null argument");
    if (nop < 1 || nop > p->cap_nop)
        return fail(HH_ERR_ARG, "hh_plan_set_operators: %d operators, the plan holds at most %d", nop, p->cap_nop);
    hh_context *c = p->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = set_device(c);
    if (rc) return rc;
    const int nc = p->dim == 2 ? 6 : 12, n = p->dim == 2 ? p->Ny : p->Nz;
    // the host staging buffers feed the previous call's copies until they
    // complete; wait for the stream before refilling them
    HIPCHK(hipStreamSynchronize(c->stream));
    const int old_real = p->real_op, old_nfx = p->ms.nfx;
    fill_table(p, nop, coef);
    if (p->real_op != old_real || p->ms.nfx != old_nfx || nop != p->nop) p->gen++;
    if ((rc = upload_modes(p))) return rc;
    const size_t cnt = (size_t)nc * n * nop;
    p->h32.resize(2 * cnt);
    for (size_t i = 0; i < cnt; i++) {
        p->h32[2 * i] = (float)p->hcoef[i].real();
        p->h32[2 * i + 1] = (float)p->hcoef[i].imag();
    }
    // stream-ordered: solves queued before this call still read the old tables
    HIPCHK(hipMemcpyAsync(p->coef64, p->hcoef.data(), cnt * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(p->coef32, p->h32.data(), cnt * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(p->coefr64, p->hr64.data(), cnt * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(p->coefr32, p->hr32.data(), cnt * 4, hipMemcpyHostToDevice, c->stream));
    p->nop = nop;
    return HH_OK;
}

extern "C" int hh_solve_host(hh_plan *p, const double *h_rhs, double *h_sol, int batch) {
    if (!p || !h_rhs || !h_sol) return fail(HH_ERR_ARG, "hh_solve_host: null argument");
    if (batch < 1) return fail(HH_ERR_ARG, "hh_solve_host: batch must be >= 1 (got %d)", batch);
    hh_context *c = p->ctx;
    int rc = set_device(c);
    if (rc) return rc;
    size_t bytes = (size_t)hh_plan_field_size(p) * batch * 16;
    void *d_rhs = nullptr, *d_sol = nullptr;
    if ((rc = dev_alloc(c, &d_rhs, bytes))) return rc;
    if ((rc = dev_alloc(c, &d_sol, bytes))) { dev_free(c, d_rhs, bytes); return rc; }
    hipError_t e = hipMemcpyAsync(d_rhs, h_rhs, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        rc = hh_solve(p, d_rhs, d_sol, batch);
        if (rc == HH_OK) {
            e = hipMemcpyAsync(h_sol, d_sol, bytes, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        }
    }
    (void)hipStreamSynchronize(c->stream);
    dev_free(c, d_rhs, bytes);
    dev_free(c, d_sol, bytes);
    if (rc) return rc;
    if (e != hipSuccess) return fail(HH_ERR_DEVICE, "host transfer failed: %s", hipGetErrorString(e));
    return HH_OK;
}

extern "C" int hh_residual(hh_plan *p, const void *d_sol, const void *d_rhs, void *d_res, int batch) {
    if (!p || !d_sol || !d_rhs || !d_res) return fail(HH_ERR_ARG, "hh_residual: null argument");
    if (batch < 1) return fail(HH_ERR_ARG, "hh_residual: batch must be >= 1");
    int rc = set_device(p->ctx);
    if (rc) return rc;
    return residual_into(p, d_sol, d_rhs, d_res, 0, batch);
}

extern "C" int hh_dst_rows(hh_context *c, int N, long nrows, const void *in, void *out) {
    if (!c || !in || !out || N < 1 || nrows < 0) return fail(HH_ERR_ARG, "hh_dst_rows: bad arguments");
    if (nrows == 0) return HH_OK;
    if (!hh::fft_supported(N) && N > 4096)
        return fail(HH_ERR_UNSUPPORTED, "DST length %d too large for the dense path", N);
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = set_device(c);
    if (rc) return rc;
    hh::Lines l = rows_lines(N, nrows);
    if (!hh::fft_supported(N) && in == out) {   // dense path is out of place
        if ((rc = ensure_scratch(c, (size_t)N * nrows * 16))) return rc;
        HIPCHK(hipMemcpyAsync(c->scratch, in, (size_t)N * nrows * 16, hipMemcpyDeviceToDevice, c->stream));
        return dst_pass(c, l, 0, 0, 0, c->scratch, out);
    }
    return dst_pass(c, l, 0, 0, 0, in, out);
}

extern "C" int hh_dst2d_layers(hh_context *c, int Nx, int Ny, long nlayers, const void *in, void *out) {
    if (!c || !in || !out || Nx < 1 || Ny < 1 || nlayers < 0) return fail(HH_ERR_ARG, "hh_dst2d_layers: bad arguments");
    if (nlayers == 0) return HH_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = set_device(c);
    if (rc) return rc;
    hh::Lines lx = rows_lines(Nx, (long)Ny * nlayers), ly = cols_lines(Nx, Ny, nlayers);
    bool fx = hh::fft_supported(Nx), fy = hh::fft_supported(Ny);
    size_t bytes = (size_t)Nx * Ny * nlayers * 16;
    if (fx && fy) {
        if ((rc = dst_pass(c, lx, 0, 0, 0, in, out))) return rc;
        return dst_pass(c, ly, 0, 0, 0, out, out);
    }
    if ((rc = ensure_scratch(c, bytes))) return rc;
    if ((rc = dst_pass(c, lx, 0, 0, 0, in, c->scratch))) return rc;
    return dst_pass(c, ly, 0, 0, 0, c->scratch, out);
}

// compact (k^2)_y (2D_Program/partial_derivatives/4th_order_first_derivative.c:4-34)
static void k2_y_compact(int Ny, double hy, const double *k2e, const double *k2yb, std::vector<cd> &k2y) {
    k2y.resize(Ny);
    for (int j = 0; j < Ny; j++) k2y[j] = 3.0 * (ld(k2e, j + 2) - ld(k2e, j)) / hy;
    k2y[0] -= ld(k2yb, 0);
    k2y[Ny - 1] -= ld(k2yb, 1);
    compact_line(k2y, 4.0);
}

extern "C" int hh_rhs2d(hh_plan *p, int order, double hx, double hy, const void *f_e, const void *u_e,
                        const void *fy_b, const void *fxx_b, const void *fyy_b, const double *k2e, const double *k2yb,
                        void *F, int batch) {
    if (!p || p->dim != 2 || !f_e || !u_e || !k2e || !F || batch < 1) return fail(HH_ERR_ARG, "hh_rhs2d: bad arguments");
    if (order != 2 && order != 4 && order != 6) return fail(HH_ERR_ARG, "order must be 2, 4 or 6 (got %d)", order);
    if (order == 6 && (!fy_b || !fxx_b || !fyy_b)) return fail(HH_ERR_ARG, "hh_rhs2d: sixth order needs boundary data");
    hh_context *c = p->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = set_device(c);
    if (rc) return rc;
    const int Nx = p->Nx, Ny = p->Ny;
    const long XY = (long)(Nx + 2) * (Ny + 2), wstride = 2 * XY + (long)Nx * Ny;
    // per-row k^2 and compact (k^2)_y, uploaded after the field work area
    size_t wbytes = order == 6 ? (size_t)wstride * batch * 16 : 0;
    const int nop = p->nop;
    size_t tbytes = (size_t)2 * Ny * nop * 16;
    if ((rc = ensure_scratch(c, wbytes + tbytes))) return rc;
    // tables [op][j]: k^2 at the interior rows, then compact (k^2)_y
    std::vector<cd> tab(2 * (size_t)Ny * nop);
    std::vector<cd> k2y;
    for (int o = 0; o < nop; o++) {
        const double *ke = k2e + 2 * (size_t)o * (Ny + 2);
        if (order == 6) k2_y_compact(Ny, hy, ke, k2yb ? k2yb + 4 * (size_t)o : nullptr, k2y);
        for (int j = 0; j < Ny; j++) {
            tab[(size_t)o * Ny + j] = ld(ke, j + 1);
            tab[(size_t)(nop + o) * Ny + j] = order == 6 ? k2y[j] : cd(0);
        }
    }
    char *tb = (char *)c->scratch + wbytes;
    HIPCHK(hipMemcpyAsync(tb, tab.data(), tbytes, hipMemcpyHostToDevice, c->stream));
    hh::Rhs2D r;
    r.order = order; r.Nx = Nx; r.Ny = Ny; r.batch = batch; r.hx = hx; r.hy = hy;
    r.f_e = f_e; r.u_e = u_e; r.fy_b = fy_b; r.fxx_b = fxx_b; r.fyy_b = fyy_b;
    r.k2c = tb; r.k2y = tb + (size_t)Ny * nop * 16; r.coef = p->coef64; r.nop = nop; r.work = c->scratch;
    TimeScope ts(c, 4, (double)Nx * Ny * batch * 32);
    hipError_t e = hh::launch_rhs2d(r, F, c->stream);
    if (e != hipSuccess) return fail(HH_ERR_DEVICE, "rhs2d launch failed: %s", hipGetErrorString(e));
    // the host table must outlive the asynchronous copy
    HIPCHK(hipStreamSynchronize(c->stream));
    return HH_OK;
}

extern "C" int hh_rhs3d(hh_plan *p, int order, double hx, double hy, double hz, const void *f_e, const void *u_e,
                        const void *fxx_f, const void *fyy_f, const void *fzz_f, const double *k2e, const double *k2d,
                        void *F, int batch) {
    if (!p || p->dim != 3 || !f_e || !u_e || !k2e || !F || batch < 1) return fail(HH_ERR_ARG, "hh_rhs3d: bad arguments");
    if (order != 2 && order != 4 && order != 6) return fail(HH_ERR_ARG, "order must be 2, 4 or 6 (got %d)", order);
    if (order == 6 && (!fxx_f || !fyy_f || !fzz_f)) return fail(HH_ERR_ARG, "hh_rhs3d: sixth order needs face data");
    hh_context *c = p->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = set_device(c);
    if (rc) return rc;
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const long E = (long)(Nx + 2) * (Ny + 2) * (Nz + 2);
    size_t wbytes = order == 6 ? (size_t)3 * E * batch * 16 : 0;
    const int nop = p->nop;
    size_t tbytes = (size_t)(2 * Nz + 2) * nop * 16;
    if ((rc = ensure_scratch(c, wbytes + tbytes))) return rc;
    // tables [op]: k^2 on the extended z grid, then compact (k^2)_z
    // (derivative_kz_4th_order, 4th_order_partials.c:296-340)
    std::vector<cd> tab((2 * (size_t)Nz + 2) * nop);
    std::vector<cd> kz(Nz);
    for (int o = 0; o < nop; o++) {
        const double *ke = k2e + 2 * (size_t)o * (Nz + 2);
        const double *kd = k2d ? k2d + 8 * (size_t)o : nullptr;
        for (int l = 0; l < Nz + 2; l++) tab[(size_t)o * (Nz + 2) + l] = ld(ke, l);
        for (int l = 0; l < Nz; l++) kz[l] = 3.0 * (ld(ke, l + 2) - ld(ke, l)) / hz;
        kz[0] -= ld(kd, 0);
        kz[Nz - 1] -= ld(kd, 1);
        compact_line(kz, 4.0);
        for (int l = 0; l < Nz; l++) tab[(size_t)nop * (Nz + 2) + (size_t)o * Nz + l] = kz[l];
    }
    char *tb = (char *)c->scratch + wbytes;
    HIPCHK(hipMemcpyAsync(tb, tab.data(), tbytes, hipMemcpyHostToDevice, c->stream));
    hh::Rhs3D r;
    r.order = order; r.Nx = Nx; r.Ny = Ny; r.Nz = Nz; r.batch = batch; r.hx = hx; r.hy = hy; r.hz = hz;
    r.f_e = f_e; r.u_e = u_e; r.fxx_f = fxx_f; r.fyy_f = fyy_f; r.fzz_f = fzz_f;
    r.k2e = tb; r.kz = tb + (size_t)(Nz + 2) * nop * 16; r.coef = p->coef64; r.nop = nop; r.work = c->scratch;
    TimeScope ts(c, 4, (double)Nx * Ny * Nz * batch * 32);
    hipError_t e = hh::launch_rhs3d(r, F, c->stream);
    if (e != hipSuccess) return fail(HH_ERR_DEVICE, "rhs3d launch failed: %s", hipGetErrorString(e));
    HIPCHK(hipStreamSynchronize(c->stream));
    return HH_OK;
}

extern "C" int hh_norms(hh_context *c, long n, const void *x, const void *y, double out[2]) {
    if (!c || !out || n < 0 || (n > 0 && !x)) return fail(HH_ERR_ARG, "hh_norms: bad arguments");
    if (n == 0) { out[0] = out[1] = 0.0; return HH_OK; }   // norms of an empty vector (x may be null)
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = set_device(c);
    if (rc) return rc;
    return norms_sync(c, n, x, 0, y, out);
}

extern "C" int hh_norms_l2(hh_context *c, long n, const void *x, const void *y, double out[2]) {
    if (!c || !out || n < 0 || (n > 0 && !x)) return fail(HH_ERR_ARG, "hh_norms_l2: bad arguments");
    if (n == 0) { out[0] = out[1] = 0.0; return HH_OK; }   // norms of an empty vector (x may be null)
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = set_device(c);
    if (rc) return rc;
    return norms_sync(c, n, x, 0, y, out, true);
}
