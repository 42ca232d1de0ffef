// Serving loop: continuous batching of mixed 2D/3D solve requests.
//
// No reference counterpart (the reference's main() solves one problem per
// process).  Built only on the public C ABI of hh_solver.h.
//
// Requests (one right-hand side each, host buffers) go to one FIFO queue.
// L lanes run concurrently, each a host thread with its own context and
// non-blocking stream, so batches on different lanes overlap on the GPU.
// An idle lane takes the oldest waiting request and every other waiting
// request of the same grid and precision (up to max_batch), provided the
// batch's device footprint fits in what the HBM budget has left; otherwise
// it waits until a running batch finishes and returns its bytes.  Requests
// arriving while batches run join the next batch: admission is continuous,
// not per wave.
//
// Each lane keeps, per grid and precision, one device-resident multi-
// operator plan with max_batch operator slots (eigenvalues and twiddle
// tables stay resident) plus right-hand-side and solution buffers; a batch
// only uploads its operators' coefficient tables (hh_plan_set_operators),
// its right-hand sides, solves, and copies the solutions back.  Resident
// plans and buffers are charged to the budget and the least recently used
// ones of an idle lane are released when a batch would not fit otherwise.
//
// A result does not depend on the batch it ran in: FP64 solves give the
// same bits alone and in any mix (the kernels treat right-hand sides
// independently and the block shapes depend on the grid only); mixed-
// precision solves stop refining on the batch's largest residual, so a
// result in a mix may take more sweeps than alone, never fewer, and agrees
// with the alone result to the refinement tolerance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <complex>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "hh_solver.h"

namespace {

using Clock = std::chrono::steady_clock;
using cd = std::complex<double>;

struct Shape {
    int dim, Nx, Ny, Nz, prec;
    bool operator<(const Shape &o) const {
        return std::tie(dim, Nx, Ny, Nz, prec) < std::tie(o.dim, o.Nx, o.Ny, o.Nz, o.prec);
    }
    bool operator==(const Shape &o) const {
        return dim == o.dim && Nx == o.Nx && Ny == o.Ny && Nz == o.Nz && prec == o.prec;
    }
    long field() const { return (long)Nx * Ny * (dim == 3 ? Nz : 1); }
    int sys() const { return dim == 2 ? Ny : Nz; }
};

struct Req {
    long id;
    Shape shape;
    hh_request r;
    Clock::time_point t_arrive, t_start, t_done;
    int status = HH_OK;
    int sweeps = 0;
    double rel = 0;
    bool done = false;
};

// Resident state of one grid on one lane.
struct Resident {
    hh_plan *plan = nullptr;
    void *d_rhs = nullptr, *d_sol = nullptr;
    int cap = 0;          // right-hand sides the buffers hold
    size_t bytes = 0;     // charged to the budget
    long last_use = 0;
};

}  // namespace

struct hh_server {
    int device = 0, lanes = 1, max_batch = 8;
    size_t budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> queue;
    std::map<long, std::unique_ptr<Req>> reqs;
    long next_id = 1;
    size_t reserved = 0;   // bytes of running batches plus resident plans
    bool stop = false;
    long clock = 0;
    std::vector<std::thread> threads;
    std::atomic<long> batches{0}, solved{0};
    std::atomic<int> lane_error{HH_OK};
    std::vector<std::map<Shape, Resident>> resident;   // per lane
    std::vector<char> busy;                             // lane is running a batch
    std::vector<size_t> lane_ws;                        // workspace charged to each lane
};

namespace {

size_t field_bytes(const Shape &s) { return (size_t)s.field() * 16; }

// Coefficient vectors of one request's operator (constant k, as the
// reference's coefficients() with k^2 sampled on the extended grid).
int request_coef(const hh_request &r, std::vector<double> &out) {
    const cd k = cd(r.k_re, r.k_im), k2 = k * k;
    if (r.dim == 2) {
        std::vector<double> k2e(2 * (size_t)(r.Ny + 2));
        for (int j = 0; j < r.Ny + 2; j++) { k2e[2 * j] = k2.real(); k2e[2 * j + 1] = k2.imag(); }
        out.assign(2 * 6 * (size_t)r.Ny, 0.0);
        double *v[6];
        for (int q = 0; q < 6; q++) v[q] = out.data() + 2 * (size_t)q * r.Ny;
        return hh_coefficients2d(r.order, r.Nx, r.Ny, r.hx, r.hy, k2e.data(), nullptr, nullptr, v[0], v[1], v[2],
                                 v[3], v[4], v[5]);
    }
    std::vector<double> k2e(2 * (size_t)(r.Nz + 2));
    for (int j = 0; j < r.Nz + 2; j++) { k2e[2 * j] = k2.real(); k2e[2 * j + 1] = k2.imag(); }
    out.assign(2 * 12 * (size_t)r.Nz, 0.0);
    double *v[12];
    for (int q = 0; q < 12; q++) v[q] = out.data() + 2 * (size_t)q * r.Nz;
    return hh_coefficients3d(r.order, r.Nx, r.Ny, r.Nz, r.hx, r.hy, r.hz, k2e.data(), nullptr, v);
}

// Device bytes a resident grid needs for cap right-hand sides: its buffers
// and the coefficient tables and eigenvalues of max_batch operators.
size_t resident_bytes(const Shape &s, int cap, int max_batch) {
    // coefficient tables per entry: complex FP64 and FP32 (16 + 8 B) and
    // their real parts (8 + 4 B) for the real-band kernels
    const size_t tables = (size_t)max_batch * (s.dim == 2 ? 6 : 12) * s.sys() * 36 + 8 * (s.Nx + s.Ny);
    return 2 * (size_t)cap * field_bytes(s) + tables;
}

// Workspace of one solve of cap right-hand sides on the server's plans
// (row-major spectra), as hh_plan_workspace_bytes states it: the dense
// DST's ping-pong field (16 B/DOF) unless every transformed length has N+1
// a power of two up to 4096 (the FFT path), and the FP32 correction of
// mixed precision (8 B/DOF).  It lives in the lane context's scratch,
// which keeps the largest size any batch needed until the lane ends, so it
// is charged to the lane, not to a resident grid.
size_t solve_ws_bytes(const Shape &s, int cap) {
    auto fft = [](int N) { const int m = N + 1; return m >= 2 && m <= 4096 && (m & (m - 1)) == 0; };
    const bool all_fft = fft(s.Nx) && (s.dim == 2 || fft(s.Ny));
    const size_t f = (size_t)cap * s.field();
    return (all_fft ? 0 : 16 * f) + (s.prec == HH_PREC_MIXED ? 8 * f : 0);
}

void release(hh_server *S, Resident &R) {
    if (R.plan) hh_plan_destroy(R.plan);
    if (R.d_rhs) (void)hipFree(R.d_rhs);
    if (R.d_sol) (void)hipFree(R.d_sol);
    S->reserved -= std::min(S->reserved, R.bytes);
    R = Resident();
}

void lane_main(hh_server *S, int lane) {
    (void)hipSetDevice(S->device);
    hh_context *ctx = nullptr;
    if (hh_context_create(S->device, &ctx) != HH_OK) {
        S->lane_error = HH_ERR_DEVICE;
        std::lock_guard<std::mutex> lk(S->mu);
        S->cv_done.notify_all();
        return;
    }
    auto &cache = S->resident[lane];
    std::vector<double> coef, one;
    for (;;) {
        std::vector<Req *> batch;
        Shape sh{};
        size_t grow = 0, wgrow = 0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop && S->queue.empty()) {
                    for (auto &kv : cache) release(S, kv.second);
                    S->reserved -= std::min(S->reserved, S->lane_ws[lane]);
                    S->lane_ws[lane] = 0;
                    lk.unlock();
                    hh_context_destroy(ctx);
                    return;
                }
                if (!S->queue.empty()) {
                    sh = S->queue.front()->shape;
                    // the oldest request and the waiting ones of its grid
                    int nb = 0;
                    for (Req *q : S->queue)
                        if (q->shape == sh && nb < S->max_batch) nb++;
                    Resident &R = cache[sh];
                    const size_t need = resident_bytes(sh, std::max(nb, R.cap), S->max_batch);
                    grow = need > R.bytes ? need - R.bytes : 0;
                    const size_t wneed = solve_ws_bytes(sh, nb);
                    wgrow = wneed > S->lane_ws[lane] ? wneed - S->lane_ws[lane] : 0;
                    // make room: drop the least recently used resident grids
                    // of idle lanes (this one included, except the grid it
                    // is about to use)
                    while (S->budget && S->reserved + grow + wgrow > S->budget) {
                        Resident *victim = nullptr;
                        for (int l = 0; l < S->lanes; l++) {
                            if (l != lane && S->busy[l]) continue;
                            for (auto &kv : S->resident[l])
                                if (kv.second.plan && !(l == lane && kv.first == sh) &&
                                    (!victim || kv.second.last_use < victim->last_use))
                                    victim = &kv.second;
                        }
                        if (!victim) break;
                        release(S, *victim);
                    }
                    if (!S->budget || S->reserved + grow + wgrow <= S->budget) {
                        for (auto it = S->queue.begin(); it != S->queue.end() && (int)batch.size() < nb;) {
                            if ((*it)->shape == sh) { batch.push_back(*it); it = S->queue.erase(it); }
                            else ++it;
                        }
                        S->reserved += grow + wgrow;
                        S->lane_ws[lane] += wgrow;
                        S->busy[lane] = 1;
                        R.last_use = ++S->clock;
                        break;
                    }
                    bool any_busy = false;
                    for (int l = 0; l < S->lanes; l++) any_busy = any_busy || S->busy[l];
                    if (!any_busy) {
                        // nothing runs and nothing more can be freed: the
                        // oldest request does not fit in the budget at all
                        Req *q = S->queue.front();
                        S->queue.pop_front();
                        q->status = HH_ERR_NOMEM;
                        q->done = true;
                        q->t_start = q->t_done = Clock::now();
                        S->cv_done.notify_all();
                        continue;
                    }
                }
                S->cv_work.wait(lk);
            }
        }
        const Clock::time_point t0 = Clock::now();
        for (Req *q : batch) q->t_start = t0;
        int nb = (int)batch.size();
        Resident &R = cache[sh];
        int rc = HH_OK;
        // grow the resident grid to nb right-hand sides
        if (!R.plan) {
            rc = request_coef(batch[0]->r, one);
            std::vector<double> ops;
            for (int i = 0; rc == HH_OK && i < S->max_batch; i++) ops.insert(ops.end(), one.begin(), one.end());
            if (rc == HH_OK)
                rc = sh.dim == 2 ? hh_plan2d_create_multi(ctx, sh.Nx, sh.Ny, S->max_batch, ops.data(), sh.prec, &R.plan)
                                 : hh_plan3d_create_multi(ctx, sh.Nx, sh.Ny, sh.Nz, S->max_batch, ops.data(), sh.prec,
                                                          &R.plan);
        }
        if (rc == HH_OK && R.cap < nb) {
            if (R.d_rhs) (void)hipFree(R.d_rhs);
            if (R.d_sol) (void)hipFree(R.d_sol);
            R.d_rhs = R.d_sol = nullptr;
            if (hipMalloc(&R.d_rhs, nb * field_bytes(sh)) != hipSuccess ||
                hipMalloc(&R.d_sol, nb * field_bytes(sh)) != hipSuccess)
                rc = HH_ERR_NOMEM;
            else
                R.cap = nb;
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            R.bytes += grow;
        }
        // operators of the batch, right-hand sides, solve, solutions back
        if (rc == HH_OK) {
            coef.clear();
            for (Req *q : batch) {
                if ((rc = request_coef(q->r, one))) break;
                coef.insert(coef.end(), one.begin(), one.end());
            }
        }
        if (rc == HH_OK) rc = hh_plan_set_operators(R.plan, nb, coef.data());
        // a request whose operator is singular to working precision (k^2 at
        // an eigenvalue) fails alone: the plan's classification of the
        // batch's operators names it, the others are solved without it
        std::vector<Req *> refused;
        if (rc == HH_OK) {
            const size_t per = coef.size() / nb;
            std::vector<Req *> keep;
            std::vector<double> kc;
            for (int b = 0; b < nb; b++) {
                int sg = 0;
                hh_plan_mode_info(R.plan, b, nullptr, &sg, nullptr);
                if (sg) {
                    refused.push_back(batch[b]);
                } else {
                    keep.push_back(batch[b]);
                    kc.insert(kc.end(), coef.begin() + b * per, coef.begin() + (b + 1) * per);
                }
            }
            if (!refused.empty()) {
                batch.swap(keep);
                nb = (int)batch.size();
                if (nb > 0) rc = hh_plan_set_operators(R.plan, nb, kc.data());
            }
        }
        hipStream_t st = (hipStream_t)hh_context_get_stream(ctx);
        for (int b = 0; rc == HH_OK && b < nb; b++)
            if (hipMemcpyAsync((char *)R.d_rhs + b * field_bytes(sh), batch[b]->r.h_rhs, field_bytes(sh),
                               hipMemcpyHostToDevice, st) != hipSuccess)
                rc = HH_ERR_DEVICE;
        if (rc == HH_OK && nb > 0) rc = hh_solve(R.plan, R.d_rhs, R.d_sol, nb);
        for (int b = 0; rc == HH_OK && b < nb; b++)
            if (hipMemcpyAsync(batch[b]->r.h_sol, (char *)R.d_sol + b * field_bytes(sh), field_bytes(sh),
                               hipMemcpyDeviceToHost, st) != hipSuccess)
                rc = HH_ERR_DEVICE;
        if (hipStreamSynchronize(st) != hipSuccess && rc == HH_OK) rc = HH_ERR_DEVICE;
        int sweeps = 0;
        double rel = 0;
        if (rc == HH_OK && sh.prec == HH_PREC_MIXED) hh_plan_last_refine(R.plan, &sweeps, &rel);
        const Clock::time_point t1 = Clock::now();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (Req *q : batch) {
                q->status = rc;
                q->sweeps = sweeps;
                q->rel = rel;
                q->t_done = t1;
                q->done = true;
            }
            for (Req *q : refused) {
                q->status = HH_ERR_SINGULAR;
                q->t_done = t1;
                q->done = true;
            }
            R.last_use = ++S->clock;
            // a failed batch leaves no half-built grid behind: its bytes go
            // back to the budget (eviction only finds grids with a plan)
            if (rc != HH_OK) release(S, R);
            S->busy[lane] = 0;
            S->batches++;
            if (rc == HH_OK) S->solved += nb;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

}  // namespace

extern "C" int hh_server_create(int device, size_t hbm_budget, int lanes, int max_batch, hh_server **out) {
    if (!out || lanes < 1 || lanes > 16 || max_batch < 1 || max_batch > 1024) return HH_ERR_ARG;
    hh_server *S = new hh_server();
    S->device = device;
    S->budget = hbm_budget;
    S->lanes = lanes;
    S->max_batch = max_batch;
    S->resident.resize(lanes);
    S->busy.assign(lanes, 0);
    S->lane_ws.assign(lanes, 0);
    for (int l = 0; l < lanes; l++) S->threads.emplace_back(lane_main, S, l);
    *out = S;
    return HH_OK;
}

extern "C" int hh_server_submit(hh_server *S, const hh_request *r, long *ticket) {
    if (!S || !r || !ticket || !r->h_rhs || !r->h_sol) return HH_ERR_ARG;
    if ((r->dim != 2 && r->dim != 3) || r->Nx < 2 || r->Ny < 2 || (r->dim == 3 && r->Nz < 2)) return HH_ERR_ARG;
    if (r->order != 2 && r->order != 4 && r->order != 6) return HH_ERR_ARG;
    if (r->precision != HH_PREC_F64 && r->precision != HH_PREC_MIXED) return HH_ERR_ARG;
    if (S->lane_error != HH_OK) return S->lane_error;
    auto q = std::make_unique<Req>();
    q->r = *r;
    q->shape = Shape{r->dim, r->Nx, r->Ny, r->dim == 3 ? r->Nz : 1, r->precision};
    q->t_arrive = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return HH_ERR_ARG;
    q->id = S->next_id++;
    *ticket = q->id;
    S->queue.push_back(q.get());
    S->reqs[q->id] = std::move(q);
    S->cv_work.notify_all();
    return HH_OK;
}

extern "C" int hh_server_wait(hh_server *S, long ticket, hh_result *res) {
    if (!S) return HH_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return HH_ERR_ARG;
    Req *q = it->second.get();
    S->cv_done.wait(lk, [&] { return q->done || S->lane_error != HH_OK; });
    if (!q->done) return S->lane_error;
    if (res) {
        res->status = q->status;
        res->queue_s = std::chrono::duration<double>(q->t_start - q->t_arrive).count();
        res->latency_s = std::chrono::duration<double>(q->t_done - q->t_arrive).count();
        res->sweeps = q->sweeps;
        res->rel_residual = q->rel;
    }
    const int st = q->status;
    S->reqs.erase(it);
    return st;
}

extern "C" int hh_server_stats(hh_server *S, long *batches, long *solved, size_t *reserved_bytes) {
    if (!S) return HH_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    if (batches) *batches = S->batches;
    if (solved) *solved = S->solved;
    if (reserved_bytes) *reserved_bytes = S->reserved;
    return HH_OK;
}

extern "C" int hh_server_destroy(hh_server *S) {
    if (!S) return HH_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &t : S->threads) t.join();
    delete S;
    return HH_OK;
}
