// Host-pointer variants of the C ABI (include/hh_solver.h): stage host
// arrays through device buffers, run the device entry point on the
// context's stream, copy back and synchronise.  These serve C callers that
// hold plain host arrays, as the reference's System does.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/hh_solver.h"

namespace {

// Device staging buffers freed on scope exit.
struct Stage {
    std::vector<void *> bufs;
    ~Stage() {
        for (void *p : bufs) (void)hipFree(p);
    }
    void *put(const double *h, size_t bytes, hipStream_t st, int *rc) {
        if (*rc) return nullptr;
        void *d = nullptr;
        if (hipMalloc(&d, bytes ? bytes : 16) != hipSuccess) { *rc = HH_ERR_NOMEM; return nullptr; }
        bufs.push_back(d);
        if (h && hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st) != hipSuccess) *rc = HH_ERR_DEVICE;
        return d;
    }
};

int finish(int rc, void *d, double *h, size_t bytes, hipStream_t st) {
    if (rc) return rc;
    if (hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st) != hipSuccess) return HH_ERR_DEVICE;
    if (hipStreamSynchronize(st) != hipSuccess) return HH_ERR_DEVICE;
    return HH_OK;
}

hipStream_t stream_of(hh_context *c) { return (hipStream_t)hh_context_get_stream(c); }

}  // namespace

extern "C" int hh_residual_host(hh_plan *p, const double *sol, const double *rhs, double *res, int batch) {
    if (!p || !sol || !rhs || !res || batch < 1) return HH_ERR_ARG;
    hipStream_t st = stream_of(hh_plan_context(p));
    size_t bytes = (size_t)hh_plan_field_size(p) * batch * 16;
    Stage s;
    int rc = 0;
    void *du = s.put(sol, bytes, st, &rc), *dF = s.put(rhs, bytes, st, &rc), *dr = s.put(nullptr, bytes, st, &rc);
    if (!rc) rc = hh_residual(p, du, dF, dr, batch);
    return finish(rc, dr, res, bytes, st);
}

extern "C" int hh_dst_rows_host(hh_context *c, int N, long nrows, const double *in, double *out) {
    if (!c || !in || !out || N < 1 || nrows < 1) return HH_ERR_ARG;
    hipStream_t st = stream_of(c);
    size_t bytes = (size_t)N * nrows * 16;
    Stage s;
    int rc = 0;
    void *d = s.put(in, bytes, st, &rc);
    if (!rc) rc = hh_dst_rows(c, N, nrows, d, d);
    return finish(rc, d, out, bytes, st);
}

extern "C" int hh_dst2d_layers_host(hh_context *c, int Nx, int Ny, long nlayers, const double *in, double *out) {
    if (!c || !in || !out || Nx < 1 || Ny < 1 || nlayers < 1) return HH_ERR_ARG;
    hipStream_t st = stream_of(c);
    size_t bytes = (size_t)Nx * Ny * nlayers * 16;
    Stage s;
    int rc = 0;
    void *di = s.put(in, bytes, st, &rc), *dout = s.put(nullptr, bytes, st, &rc);
    if (!rc) rc = hh_dst2d_layers(c, Nx, Ny, nlayers, di, dout);
    return finish(rc, dout, out, bytes, st);
}

extern "C" int hh_rhs2d_host(hh_plan *p, int order, double hx, double hy, const double *f_e, const double *u_e,
                             const double *fy_b, const double *fxx_b, const double *fyy_b, const double *k2e,
                             const double *k2yb, double *F) {
    int dims[3];
    if (!p || hh_plan_dims(p, dims) != 2 || !f_e || !u_e || !F) return HH_ERR_ARG;
    const int Nx = dims[0], Ny = dims[1];
    hipStream_t st = stream_of(hh_plan_context(p));
    size_t E = (size_t)(Nx + 2) * (Ny + 2) * 16, R = (size_t)(2 * (Nx + 2) + 2 * (Ny + 2)) * 16;
    Stage s;
    int rc = 0;
    void *df = s.put(f_e, E, st, &rc), *du = s.put(u_e, E, st, &rc);
    void *dfy = fy_b ? s.put(fy_b, (size_t)2 * Nx * 16, st, &rc) : nullptr;
    void *dfxx = fxx_b ? s.put(fxx_b, R, st, &rc) : nullptr;
    void *dfyy = fyy_b ? s.put(fyy_b, R, st, &rc) : nullptr;
    void *dF = s.put(nullptr, (size_t)Nx * Ny * 16, st, &rc);
    if (!rc) rc = hh_rhs2d(p, order, hx, hy, df, du, dfy, dfxx, dfyy, k2e, k2yb, dF, 1);
    return finish(rc, dF, F, (size_t)Nx * Ny * 16, st);
}

extern "C" int hh_rhs3d_host(hh_plan *p, int order, double hx, double hy, double hz, const double *f_e,
                             const double *u_e, const double *fxx_f, const double *fyy_f, const double *fzz_f,
                             const double *k2e, const double *k2d, double *F) {
    int dims[3];
    if (!p || hh_plan_dims(p, dims) != 3 || !f_e || !u_e || !F) return HH_ERR_ARG;
    const size_t X = dims[0] + 2, Y = dims[1] + 2, Z = dims[2] + 2;
    hipStream_t st = stream_of(hh_plan_context(p));
    Stage s;
    int rc = 0;
    void *df = s.put(f_e, X * Y * Z * 16, st, &rc), *du = s.put(u_e, X * Y * Z * 16, st, &rc);
    void *dx = fxx_f ? s.put(fxx_f, 2 * Y * Z * 16, st, &rc) : nullptr;
    void *dy = fyy_f ? s.put(fyy_f, 2 * X * Z * 16, st, &rc) : nullptr;
    void *dz = fzz_f ? s.put(fzz_f, 2 * X * Y * 16, st, &rc) : nullptr;
    size_t nF = (size_t)dims[0] * dims[1] * dims[2] * 16;
    void *dF = s.put(nullptr, nF, st, &rc);
    if (!rc) rc = hh_rhs3d(p, order, hx, hy, hz, df, du, dx, dy, dz, k2e, k2d, dF, 1);
    return finish(rc, dF, F, nF, st);
}
