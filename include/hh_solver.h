/*
 * hh_solver.h -- C ABI of the MI355X compact-Helmholtz direct solver.
 *
 * Plain pointers and sizes only.  Complex numbers are interleaved
 * (re, im) doubles: the bytes of C99 `double _Complex`, fftw_complex and
 * hipDoubleComplex.  Fields are row-major, index i + j*Nx (+ l*Nx*Ny), the
 * layout of the reference's System arrays (2D_Program/headers/structs.h:9-30,
 * 3D_Program/headers/structs.h:11-33).  A batch of right-hand sides is
 * stored back to back (stride Nx*Ny or Nx*Ny*Nz).
 *
 * Every function returns HH_OK (0) or a status code; hh_last_error()
 * gives the message of the last failure on the calling thread.  Functions
 * taking `d_` pointers expect device memory on the context's device and run
 * asynchronously on the context's stream; `h_` pointers are host memory and
 * the call returns when the result is in host memory.
 */
#ifndef HH_SOLVER_H
#define HH_SOLVER_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HH_OK = 0,
    HH_ERR_ARG = 1,          /* invalid size, order, pointer or batch */
    HH_ERR_DEVICE = 2,       /* HIP runtime failure */
    HH_ERR_NOMEM = 3,        /* device allocation failed or over budget */
    HH_ERR_UNSUPPORTED = 4,  /* shape outside the kernels' limits */
    HH_ERR_SINGULAR = 5      /* an operator is singular to working precision (resonance) */
};

enum { HH_PREC_F64 = 0, HH_PREC_MIXED = 1 };

typedef struct hh_context hh_context;
typedef struct hh_plan hh_plan;

const char *hh_last_error(void);
const char *hh_status_string(int status);

/* ---- context: device, stream, twiddle/DST tables, scratch ---- */
int hh_context_create(int device, hh_context **ctx);
int hh_context_destroy(hh_context *ctx);
/* run on a caller stream (hipStream_t); NULL selects the HIP default (null)
 * stream.  A new context runs on a private non-blocking stream. */
int hh_context_set_stream(hh_context *ctx, void *stream);
void *hh_context_get_stream(hh_context *ctx);
int hh_context_synchronize(hh_context *ctx);
/* limit on device bytes the context may allocate (0 = no limit) */
int hh_context_set_memory_budget(hh_context *ctx, size_t bytes);
size_t hh_context_bytes_in_use(const hh_context *ctx);
/* Kernel timing with HIP events recorded on the launch stream around each
 * pass.  Categories: 0 = DST (FFT), 1 = DST (dense), 2 = tridiagonal,
 * 3 = residual / refinement update, 4 = right-hand side.  read_timing
 * synchronises the stream and returns, per category, the summed kernel
 * milliseconds, the number of launches and the algorithmic bytes moved
 * (each element read once and written once). */
int hh_context_enable_timing(hh_context *ctx, int on);
/* DST-I of power-of-two N+1.  Bit 0 of algo: 0 = (N+1)-point FFT of the
 * symmetrised row plus a prefix-sum post-pass, 1 = 2(N+1)-point FFT of the
 * odd extension; bit 1: 64-thread blocks instead of 256; bit 2: for the
 * (N+1)-point transform of strided lines with 128 <= N+1 <= 256, 512-thread
 * blocks (tiles of twice the lines; same bits).  All are
 * orthonormal DST-I to roundoff; the prefix sums of the first make its error
 * grow like sqrt(N) rather than log N (7e-14 relative for a transform and
 * its inverse at N+1 = 4096, against 5e-16).  -1 (default): the
 * (N+1)-point transform, in 64-thread blocks for contiguous FP64 rows with
 * N+1 >= 512, 512-thread blocks for strided lines in FP32 at 128 <= N+1 <=
 * 256, and 256-thread blocks otherwise.  Other N+1 take the dense DST (the
 * orthonormal sine matrix times the lines) on the matrix cores (FP64 or
 * FP32 MFMA); bit 3 of algo selects the vector-FMA form of it instead. */
int hh_context_set_dst_algorithm(hh_context *ctx, int algo);
int hh_context_read_timing(hh_context *ctx, double ms[5], long launches[5], double bytes[5]);
int hh_context_reset_timing(hh_context *ctx);

/* ---- stencil coefficients (host, O(N)) ----
 * 2D: replaces coefficients() of 2D_Program/main_functions/coefficients.c:
 * 9-67.  k2e: k^2 at y0 + j*hy, j = 0..Ny+1 (complex, 2*(Ny+2) doubles);
 * k2y, k2yy: analytic (k^2)_y and (k^2)_yy at the interior rows (complex,
 * 2*Ny doubles, NULL = zero; sixth order only).  Outputs: the reference's
 * bp, ap, b, a, bm, am (each complex, length Ny). */
int hh_coefficients2d(int order, int Nx, int Ny, double hx, double hy, const double *k2e,
                      const double *k2y, const double *k2yy, double *bp, double *ap, double *b,
                      double *a, double *bm, double *am);
/* 3D: replaces coefficients() of 3D_Program/main_functions/coefficients.c:
 * 7-111 (with derivative_kz/kzz_4th_order of 4th_order_partials.c:296-385).
 * k2e: k^2 at z0 + l*hz, l = 0..Nz+1; k2d: analytic (k^2)_z(z0), (k^2)_z(z1),
 * (k^2)_zz(z0), (k^2)_zz(z1) (complex, 8 doubles; NULL = zero).
 * coef[12]: a, b, c, d, ap, bp, cp, dp, am, bm, cm, dm (each length Nz). */
int hh_coefficients3d(int order, int Nx, int Ny, int Nz, double hx, double hy, double hz,
                      const double *k2e, const double *k2d, double *const coef[12]);

/* ---- Robin / Sommerfeld closure of the tridiagonal axis (host, O(N) / O(N^2)) ----
 * No counterpart in the reference, which solves Dirichlet problems only
 * (2D rhs.c:126-148, 3D rhs.c:167-174); paper.pdf p.2 eq.(2) names such
 * boundaries.  The first (lo) and/or last (hi) row of the tridiagonal axis
 * (y in 2D, z in 3D) lies on the boundary, where du/dn + alpha u = g (n the
 * outward normal; alpha = -i k gives the first-order Sommerfeld condition).
 * The ghost row of the central difference is eliminated into the band
 * vectors, in place: coef as produced by hh_coefficients2d/3d (2D: bp, ap,
 * b, a, bm, am; 3D: the twelve vectors).  corr_lo / corr_hi (Nx or Nx*Ny
 * complex, NULL = not wanted) receive the term to add to that boundary row
 * of the right-hand side, -2h (ghost-row stencil) g.  alpha_* (one complex)
 * NULL = that side stays Dirichlet; g_* NULL = zero data.  The closure is
 * second-order accurate whatever the interior scheme.  Needs Ny (Nz) >= 2. */
int hh_robin2d(int Nx, int Ny, double h, const double *alpha_lo, const double *g_lo,
               const double *alpha_hi, const double *g_hi, double *const coef[6], double *corr_lo,
               double *corr_hi);
int hh_robin3d(int Nx, int Ny, int Nz, double h, const double *alpha_lo, const double *g_lo,
               const double *alpha_hi, const double *g_hi, double *const coef[12], double *corr_lo,
               double *corr_hi);

/* ---- plans: one operator, device-resident coefficients and eigenvalues ----
 * Replace the precomputation of LU() (2D_Program/main_functions/LU.c:3-46,
 * 3D_Program/main_functions/LU.c:4-81): no N^2/N^3 factors are stored, the
 * per-mode tridiagonal bands are rebuilt on the fly from these vectors. */
int hh_plan2d_create(hh_context *ctx, int Nx, int Ny, const double *bp, const double *ap,
                     const double *b, const double *a, const double *bm, const double *am,
                     int precision, hh_plan **plan);
int hh_plan3d_create(hh_context *ctx, int Nx, int Ny, int Nz, const double *const coef[12],
                     int precision, hh_plan **plan);
/* One operator per problem of a batch (no reference counterpart: the
 * reference solves one problem per call).  coef holds nop operators back to
 * back, each as the six (2D: bp, ap, b, a, bm, am, length Ny) or twelve (3D:
 * a..dm, length Nz) complex vectors; right-hand side b of a solve uses
 * operator b % nop.  All operators share the grid. */
int hh_plan2d_create_multi(hh_context *ctx, int Nx, int Ny, int nop, const double *coef, int precision,
                           hh_plan **plan);
int hh_plan3d_create_multi(hh_context *ctx, int Nx, int Ny, int Nz, int nop, const double *coef, int precision,
                           hh_plan **plan);
int hh_plan_num_operators(const hh_plan *plan);
int hh_plan_destroy(hh_plan *plan);
hh_context *hh_plan_context(const hh_plan *plan);
/* grid sizes of a plan: dims[0..2] = Nx, Ny, Nz (Nz = 1 in 2D); returns 2 or 3 */
int hh_plan_dims(const hh_plan *plan, int dims[3]);
/* elements of one right-hand side */
long hh_plan_field_size(const hh_plan *plan);
/* device bytes one solve of `batch` right-hand sides needs beyond its
 * input and output (scratch for non power-of-two sizes, refinement) */
size_t hh_plan_workspace_bytes(const hh_plan *plan, int batch);

/* ---- solve: replaces solver(System) ----
 * 2D_Program/solver_functions/solver.c:14-117 and
 * 3D_Program/solver_functions/solver.c:55-249 / :263-413.
 * sol = A^{-1} rhs for `batch` right-hand sides; d_sol may equal d_rhs.
 * HH_PREC_MIXED plans solve in FP32 and refine with the FP64 compact
 * residual until ||r||/||rhs|| <= tol or max_iter sweeps (hh_plan_set_refine). */
int hh_solve(hh_plan *plan, const void *d_rhs, void *d_sol, int batch);
int hh_solve_host(hh_plan *plan, const double *h_rhs, double *h_sol, int batch);
/* hh_solve that allocates nothing.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * results as hh_solve; the workspace (hh_plan_workspace_bytes) comes from a
 * static array of 128 MiB per device, held until the process ends, instead
 * of the context's scratch.  A solve needing more returns
 * HH_ERR_UNSUPPORTED; a plan table missing (the context's DST algorithm
 * changed after the plan was created) returns HH_ERR_UNSUPPORTED too.  Every
 * call on one device shares that array: calls that overlap on one GPU (two
 * contexts or streams of one device running at the same time) race. */
int hh_solve_noalloc(hh_plan *plan, const void *d_rhs, void *d_sol, int batch);
/* tol > 0: refine until ||r||inf/||rhs||inf <= tol or max_iter sweeps (each
 * check reads two norms back to the host).  tol = 0: exactly max_iter sweeps
 * with no norm and no host synchronisation, the form a HIP graph can hold. */
int hh_plan_set_refine(hh_plan *plan, double tol, int max_iter);
/* sweeps and final relative residual of the last mixed solve (-1 when the
 * sweeps were fixed and no residual was measured) */
int hh_plan_last_refine(const hh_plan *plan, int *sweeps, double *rel_residual);
/* Replace the operators of a plan made by hh_plan*_create_multi, with
 * 1 <= nop <= the count it was created with (no reference counterpart: lets
 * a serving loop reuse one device-resident plan per grid for every batch).
 * coef has the layout of the create_multi call.  Waits for the work already
 * queued on the context's stream, then queues the table copies on it. */
int hh_plan_set_operators(hh_plan *plan, int nop, const double *coef);
/* Near-resonant modes (no reference counterpart; the reference's LU() divides
 * by whatever pivot it meets, 2D_Program/main_functions/LU.c:36,
 * 3D_Program/main_functions/LU.c:68).  For every operator and mode the plan
 * bounds the condition number of the mode's tridiagonal system by Varah's
 * bound max_q(|sub|+|dia|+|sup|) / min_q(|dia|-|sub|-|sup|).  Modes that are
 * not diagonally dominant or whose bound exceeds kappa (default 4096) are
 * near-resonant: their LU factors are computed on the host in the
 * reference's operation order and the device runs the reference's
 * sequential substitution on them (FP64 arithmetic, also in mixed
 * precision); the other modes use the partitioned elimination.  A
 * near-resonant mode whose estimated condition number exceeds 1e14, or
 * whose pivot is zero or not finite, makes its operator singular: hh_solve
 * (and hh_graph_create) then returns HH_ERR_SINGULAR without touching the
 * output when a right-hand side of the batch uses that operator.
 * set_resonance_threshold re-classifies the current operators (kappa = 0
 * routes every mode to the sequential substitution, a huge kappa none but
 * the indefinite ones).  mode_info: number of near-resonant modes of
 * operator op, whether it is singular, and the largest condition estimate
 * among its near-resonant modes.  flagged_modes writes up to cap of their
 * mode indices and returns their number. */
int hh_plan_set_resonance_threshold(hh_plan *plan, double kappa);
/* Host only (no device): the same classification of one operator, coef in
 * the layout of hh_plan*_create_multi with nop = 1.  Writes up to cap of the
 * near-resonant modes, their number, whether the operator is singular and
 * the largest condition estimate among its near-resonant modes. */
int hh_operator_modes(int dim, int Nx, int Ny, int Nz, const double *coef, double kappa, int *nflagged,
                      int *modes, int cap, int *singular, double *kappa_worst);
/* Host only: the reference-order factors of one mode, 3n complex values
 * (L_q, R_q, Up_q) for q = 0..n-1, L = the reference's L, Up its Up (0 on
 * the last row), R the back-substitution multiplier (2D: 1/U, 3D: the
 * reference's stored reciprocal U).  Returns HH_ERR_SINGULAR (factors still
 * written) when the operator is singular to working precision. */
int hh_operator_mode_factors(int dim, int Nx, int Ny, int Nz, const double *coef, int mode, double *factors);
int hh_plan_mode_info(const hh_plan *plan, int op, int *flagged, int *singular, double *kappa);
int hh_plan_flagged_modes(const hh_plan *plan, int op, int *modes, int cap);
/* Block shape of the tridiagonal kernel: W neighbouring modes per block,
 * rows_per_thread in {4, 8, 16} (the chunk count follows), rhs_per_thread
 * in {1, 2, 4} right-hand sides of one operator handled together by a
 * thread (the operator's pivots are computed once for them).  Set by the
 * autotuner; (0, 0, *) restores the built-in heuristic.  Shapes differ in
 * speed only: every right-hand side gives the same bits for a given W and
 * rows_per_thread, whatever else is in its batch and whatever rhs_per_thread.
 * get_tri_shape reports the shape a solve uses; for the built-in heuristic
 * it reports rhs_per_thread = 1, though a solve with real bands at 16 rows
 * per thread takes 2 when its batch gives 2048 blocks or more (in mixed
 * precision, when an operator also has two or more right-hand sides). */
int hh_plan_set_tri_shape(hh_plan *plan, int W, int rows_per_thread, int rhs_per_thread);
int hh_plan_get_tri_shape(const hh_plan *plan, int *W, int *P, int *rows_per_thread, int *rhs_per_thread);
/* 2D FFT path: layout of the spectrum between the row DSTs and the mode
 * solves.  0 = row-major, in place in d_sol; a power of two tw (a multiple
 * of the tridiagonal W) = tiles of tw neighbouring modes holding all rows
 * contiguously, in workspace (one more field per right-hand side); -1 =
 * tiles of the tridiagonal W.  Default 0.  Tiles make the mode solves
 * faster and the row DSTs slower; layouts differ in speed only. */
int hh_plan_set_tile_width(hh_plan *plan, int tw);
/* Mode solves with real bands.  When every stencil coefficient of the
 * plan's operators has a zero imaginary part (real k), the tridiagonal
 * kernels run with real band arithmetic (complex right-hand sides); this
 * is chosen again by every hh_plan_set_operators (a recorded graph whose
 * plan changes band kind refuses to replay).  mode -1 = automatic
 * (default), 0 = complex-band kernels always.  Both give the oracle's
 * solution to the same tolerance; they differ in rounding and speed.
 * No counterpart in the reference (its LU factors are complex,
 * 2D_Program/main_functions/LU.c:3-46).  is_real_operator: 1 if the next
 * solve uses the real-band kernels. */
int hh_plan_set_real_operator(hh_plan *plan, int mode);
int hh_plan_is_real_operator(const hh_plan *plan);

/* ---- graphs: a fixed list of solves replayed with one launch ----
 * No reference counterpart (the reference solves one problem per call): for
 * batches of uneven problems, e.g. one plan per grid size.  Solve i runs
 * plans[i] on batch[i] right-hand sides d_rhs[i] -> d_sol[i].  All plans
 * belong to ctx; mixed-precision plans must use fixed sweeps
 * (hh_plan_set_refine(plan, 0, n)).  Tables are built at creation and the
 * graph owns its workspace (freed by hh_graph_destroy), so a launch
 * allocates nothing and later solves on the context do not disturb it; the
 * buffers must stay valid.  Operator values changed by
 * hh_plan_set_operators reach later replays when nothing structural
 * changed; a plan whose band kind, operator count, near-resonant mode count,
 * device tables or shape settings changed since the recording makes
 * hh_graph_launch return HH_ERR_ARG (record the graph again). */
typedef struct hh_graph hh_graph;
int hh_graph_create(hh_context *ctx, int nsolves, hh_plan *const *plans, void *const *d_rhs, void *const *d_sol,
                    const int *batch, hh_graph **graph);
/* replay on the context's stream (asynchronous) */
int hh_graph_launch(hh_graph *graph);
int hh_graph_num_nodes(const hh_graph *graph);
int hh_graph_destroy(hh_graph *graph);

/* ---- residual: replaces residual(System) ----
 * 2D_Program/main_functions/residual.c:3-34 (9-point) and
 * 3D_Program/main_functions/residual.c:42-96 (27-point): res = A sol - rhs. */
int hh_residual(hh_plan *plan, const void *d_sol, const void *d_rhs, void *d_res, int batch);

/* ---- transforms ----
 * hh_dst_rows replaces forwardDST/reverseDST (2D_Program/solver_functions/
 * DST.c:6-44): orthonormal DST-I of nrows contiguous rows of length N.
 * hh_dst2d_layers replaces DST()/DST1()/DST2() (3D_Program/solver_functions/
 * DST.c:13-204): DST-I along x then y of nlayers Nx*Ny layers.
 * Out of place or in place. */
int hh_dst_rows(hh_context *ctx, int N, long nrows, const void *d_in, void *d_out);
int hh_dst2d_layers(hh_context *ctx, int Nx, int Ny, long nlayers, const void *d_in, void *d_out);

/* ---- compact right-hand side: replaces rhs(System) ----
 * 2D_Program/main_functions/rhs.c:4-150 + partial_derivatives (all files).
 * Per right-hand side (device, complex): f_e = f on the extended
 * (Nx+2)(Ny+2) grid; u_e = Dirichlet data on the extended grid; fy_b =
 * f_y at (x_i, y0) then (x_i, y1) (2*Nx); fxx_b, fyy_b = f_xx, f_yy on the
 * ghost ring: bottom row, top row, left column, right column
 * (2(Nx+2)+2(Ny+2)).  Host, complex: k2e = k^2 at y0 + j*hy (Ny+2), k2yb =
 * (k^2)_y at y0, y1 (2), each repeated per operator of a multi-operator
 * plan.  Orders 2 and 4 ignore fy_b, fxx_b, fyy_b, k2yb. */
int hh_rhs2d(hh_plan *plan, int order, double hx, double hy, const void *d_f_e, const void *d_u_e,
             const void *d_fy_b, const void *d_fxx_b, const void *d_fyy_b, const double *k2e,
             const double *k2yb, void *d_F, int batch);
/* 3D_Program/main_functions/rhs.c:13-294 + partial_derivatives.  f_e =
 * hz^2 f on the extended grid, u_e = u on the extended grid (z faces used);
 * fxx_f = hz^2 f_xx on the x faces [face][l][j] (2(Ny+2)(Nz+2)), fyy_f on
 * the y faces [face][l][i], fzz_f on the z faces [face][j][i]; host: k2e =
 * k^2 at z0 + l*hz (Nz+2), k2d as in hh_coefficients3d. */
int hh_rhs3d(hh_plan *plan, int order, double hx, double hy, double hz, const void *d_f_e,
             const void *d_u_e, const void *d_fxx_f, const void *d_fyy_f, const void *d_fzz_f,
             const double *k2e, const double *k2d, void *d_F, int batch);

/* ---- host-pointer variants (copy in, run, copy out, synchronise) ---- */
int hh_residual_host(hh_plan *plan, const double *h_sol, const double *h_rhs, double *h_res, int batch);
int hh_dst_rows_host(hh_context *ctx, int N, long nrows, const double *h_in, double *h_out);
int hh_dst2d_layers_host(hh_context *ctx, int Nx, int Ny, long nlayers, const double *h_in, double *h_out);
/* sizes: f_e, u_e (Nx+2)(Ny+2); fy_b 2Nx; rings 2(Nx+2)+2(Ny+2); F Nx*Ny */
int hh_rhs2d_host(hh_plan *plan, int order, double hx, double hy, const double *f_e, const double *u_e,
                  const double *fy_b, const double *fxx_b, const double *fyy_b, const double *k2e,
                  const double *k2yb, double *F);
int hh_rhs3d_host(hh_plan *plan, int order, double hx, double hy, double hz, const double *f_e,
                  const double *u_e, const double *fxx_f, const double *fyy_f, const double *fzz_f,
                  const double *k2e, const double *k2d, double *F);

/* ---- norms: normInf of 2D_Program/helpers_functions/norms.c:3-13 ----
 * out[0] = max |x|, out[1] = max |x - y| (0 when d_y is NULL); blocks. */
int hh_norms(hh_context *ctx, long n, const void *d_x, const void *d_y, double out[2]);
/* normL2 of 3D_Program/helpers_functions/norms.c:15-24 on the device:
 * out[0] = sqrt(sum |x|^2), out[1] = sqrt(sum |x - y|^2) (0 when d_y is
 * NULL), summed in a fixed order (the same bits from run to run); blocks. */
int hh_norms_l2(hh_context *ctx, long n, const void *d_x, const void *d_y, double out[2]);

/* ---- serving: continuous batching of mixed 2D/3D requests ----
 * No reference counterpart (the reference's main() solves one problem per
 * process).  A server runs `lanes` worker threads, each with its own context
 * and stream; an idle lane takes the oldest waiting request together with
 * the other waiting requests of the same grid and precision (at most
 * max_batch), if the batch fits in the HBM budget (0 = unlimited), and
 * solves them as one batched multi-operator solve on its resident plan for
 * that grid.  Requests arriving meanwhile join later batches.
 * One request: constant wavenumber k on the grid with spacings h*, compact
 * scheme of the given order, one right-hand side in host memory (complex
 * interleaved, row-major); the solution is written to h_sol.  Both buffers
 * must stay valid until hh_server_wait returns for the ticket.  A request
 * whose operator is singular to working precision (k^2 at a discrete
 * eigenvalue) completes with HH_ERR_SINGULAR and h_sol untouched; the
 * requests batched with it are solved without it. */
typedef struct hh_server hh_server;
typedef struct {
    int dim, Nx, Ny, Nz;      /* Nz ignored in 2D */
    int order;                /* 2, 4 or 6 */
    int precision;            /* HH_PREC_F64 or HH_PREC_MIXED */
    double k_re, k_im;
    double hx, hy, hz;
    const double *h_rhs;
    double *h_sol;
} hh_request;
typedef struct {
    int status;               /* HH_OK or the solve's error */
    double queue_s;           /* arrival to start of its batch */
    double latency_s;         /* arrival to solution in host memory */
    int sweeps;               /* refinement sweeps (mixed precision) */
    double rel_residual;      /* final relative residual of its batch (mixed) */
} hh_result;
int hh_server_create(int device, size_t hbm_budget, int lanes, int max_batch, hh_server **server);
/* non-blocking; *ticket identifies the request for hh_server_wait */
int hh_server_submit(hh_server *server, const hh_request *req, long *ticket);
/* blocks until the request is solved; returns its status (once per ticket) */
int hh_server_wait(hh_server *server, long ticket, hh_result *result);
int hh_server_stats(hh_server *server, long *batches, long *solved, size_t *reserved_bytes);
/* finishes every submitted request, then stops the lanes */
int hh_server_destroy(hh_server *server);

#ifdef __cplusplus
}
#endif
#endif
