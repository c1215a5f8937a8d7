/*
 * chidg_hip.h -- C ABI of the MI355X (gfx950) linear-solve library for ChiDG's
 * Newton-Krylov loop: block-sparse matvec, block ILU(0) factor/apply and
 * restarted flexible GMRES.
 *
 * The entry points are what an iso_c_binding shim inside ChiDG binds (see
 * chidg_amd/fortran/mod_chidg_hip.f90).  Each declaration names the reference
 * interface it replaces (paths relative to the ChiDG source tree).
 *
 * Conventions
 *   - Every function returns CDG_OK (0) or a CDG_ERR_* code; the message of the
 *     last failure on a context is cdg_last_error(ctx).
 *   - Vector / block pointers may be host memory or device memory of the
 *     context's GPU; the library detects which.  Calls are synchronous: results
 *     are complete when the call returns (the reference semantics).
 *   - Blocks are dense n x n, column-major (type_densematrix.f90:60,177),
 *     n = nterms*nfields; vectors are element-major with the reference's
 *     intra-element order (type_densevector.f90:214,327).
 *   - Types: double, int32_t (rk = real64, ik = int32, mod_kinds.f90.in).
 *   - One context per host thread; calls on one context are not reentrant.
 */
#ifndef CHIDG_HIP_H
#define CHIDG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (messenger.f90:152-310 signals FATAL; here they return) ---- */
#define CDG_OK               0
#define CDG_ERR_ARG          1   /* invalid argument / pattern                    */
#define CDG_ERR_HIP          2   /* HIP runtime error                             */
#define CDG_ERR_SINGULAR     3   /* singular diagonal block (mod_inv.f90:126-128) */
#define CDG_ERR_NO_TRANSPOSE 4   /* lblks entry without transposed partner
                                    (type_domain_matrix.f90:620)                  */
#define CDG_ERR_OOM          5   /* device allocation failed                      */
#define CDG_ERR_NAN          6   /* non-finite residual (type_newton.f90:197-200) */
#define CDG_ERR_STATE        7   /* call out of order (e.g. apply before update)  */
#define CDG_ERR_UNSUPPORTED  8   /* block size / precision not built              */

/* ---- block kinds in a pattern --------------------------------------------- */
#define CDG_KIND_LOCAL       0   /* lblks entry: interior face or diagonal        */
#define CDG_KIND_COUPLING    1   /* chi_blks / bc_blks / hb_blks: matvec only     */

/* ---- preconditioners (mod_preconditioner.f90:49-108 factory names) -------- */
#define CDG_PREC_IDENTITY    0   /* 'identity'                                    */
#define CDG_PREC_ILU0        1   /* 'ILU0'                                        */
#define CDG_PREC_JACOBI      2   /* 'jacobi' (precon_jacobi.f90:95-215): block
                                    Jacobi, Dinv_i = inv(A_ii), z_i = Dinv_i v_i  */

/* ---- storage precision of the ILU(0) factor -------------------------------- */
#define CDG_FP64             0   /* reference precision                           */
#define CDG_FP32             1   /* factor and apply in FP32                      */
#define CDG_BF16             2   /* factor in FP32, stored BF16, FP32 accumulate  */

#define CDG_CGS              0   /* orthogonalization = 'CGS'                     */
#define CDG_MGS              1   /* orthogonalization = 'MGS'                     */

typedef struct cdg_ctx  cdg_ctx;
typedef struct cdg_mat  cdg_mat;
typedef struct cdg_prec cdg_prec;

/* linear_solver_t parameters (type_linear_solver.f90:21-40, set() :113-131). */
typedef struct {
    int32_t nkrylov;                 /* restart length                        */
    double  tol;                     /* absolute tolerance on |p(j+1)|        */
    double  rtol;                    /* relative to the first cycle's ||r0||  */
    int32_t nmax;                    /* iteration cap, <= 0 unlimited         */
    int32_t orthogonalization;       /* CDG_CGS / CDG_MGS                     */
    int32_t inner_fgmres;            /* nested FGMRES correction on/off       */
    int32_t inner_nkrylov;
    double  inner_tol;
    double  inner_rtol;
    int32_t inner_nmax;
    int32_t inner_orthogonalization;
    int32_t update_preconditioner;   /* 0: refactor only when the matrix stamp
                                        differs from the factor's stamp
                                        (solver_controller_t, type_solver_controller.f90:89-124);
                                        1: always (solve called without a controller,
                                        type_fgmres.f90:118-122) */
} cdg_fgmres_params;

/* What fgmres_t%print_report exposes (type_fgmres.f90:512-542), seconds. */
typedef struct {
    int32_t niter;
    double  residual_norm;           /* ||b - A x|| after the solve           */
    int32_t nrestart;
    int32_t nreorth;                 /* re-orthogonalized Arnoldi steps        */
    int32_t factored;                /* 1 if the preconditioner was updated    */
    double  t_total;
    double  t_factor;
    double  t_precon;
    double  t_mv;
    double  t_dot;
    double  t_norm;
} cdg_fgmres_stats;

/* ---- context ---------------------------------------------------------------- */
int         cdg_ctx_create(int32_t device, cdg_ctx **out);
int         cdg_ctx_destroy(cdg_ctx *ctx);
const char *cdg_last_error(const cdg_ctx *ctx);
/* Fills the number of bytes of device memory the context currently holds. */
int         cdg_ctx_mem_used(const cdg_ctx *ctx, int64_t *bytes);
/* The context's HIP stream (hipStream_t) so callers can order their own work
 * and record events around the *_async calls.  No reference counterpart. */
int         cdg_ctx_stream(const cdg_ctx *ctx, void **stream);
/* Ordering against a caller's stream (a hipStream_t of the context's device,
 * NULL the legacy default stream), without a host wait; no reference
 * counterpart.  cdg_ctx_wait_stream: work the context queues from now on
 * starts after the work already queued on `stream` (call it before passing
 * device vectors that stream produces).  cdg_ctx_signal_stream: work queued
 * on `stream` from now on starts after the context's queued work (call it
 * after a *_async call whose outputs that stream consumes). */
int         cdg_ctx_wait_stream(cdg_ctx *ctx, void *stream);
int         cdg_ctx_signal_stream(cdg_ctx *ctx, void *stream);

/* ---- matrix: chidg_matrix_t (type_chidg_matrix.f90:38-83) --------------------
 * Pattern: per block row (element, time level) the lblks entries followed by
 * the coupling-only entries; col[b] = parent row, kind[b] = CDG_KIND_*.
 * Replaces chidg_matrix%init (type_chidg_matrix.f90:221-248) with backend 'hip'. */
int cdg_mat_create(cdg_ctx *ctx, int32_t nrows, int32_t block_n, const int32_t *row_ptr,
                   const int32_t *col, const int32_t *kind, cdg_mat **out);
int cdg_mat_destroy(cdg_ctx *ctx, cdg_mat *A);
/* Upload all nnzb blocks, pattern order, each n*n column-major; stamps the
 * matrix.  Called from chidg_matrix%assemble (type_chidg_matrix.f90:864-867). */
int cdg_mat_upload(cdg_ctx *ctx, cdg_mat *A, const double *blocks);
int cdg_mat_download(cdg_ctx *ctx, const cdg_mat *A, double *blocks);
/* Device-side synthetic values (bench/test input, no reference counterpart):
 * block (r,c) = g*s_diag (+2 on r==c) for diagonal kind-0 blocks, g*s_off
 * otherwise; g from a counter hash of (seed, block, c*n+r).  Stamps the matrix. */
int cdg_mat_fill_synthetic(cdg_ctx *ctx, cdg_mat *A, uint64_t seed, double s_diag, double s_off);
/* Synthetic vector of length nrows*n: g(seed', index)*scale (same hash). */
int cdg_vec_fill_synthetic(cdg_ctx *ctx, const cdg_mat *A, uint64_t seed, double scale, double *v);
int cdg_mat_stamp(const cdg_mat *A, uint64_t *stamp);
/* Rank of each block row, the reference's parent_proc of every block
 * (densematrix_t, type_densematrix.f90:35-94); part = NULL: one rank.  ILU(0)
 * uses a kind-0 block only when its parent row has the row's rank, the
 * reference's parent_proc == IRANK test (precon_ILU0.f90:180,222,344,368), so
 * a partitioned matrix gets the block-Jacobi ILU(0) a multi-rank run computes,
 * all ranks on one GPU.  The matvec uses every block.  Preconditioners created
 * afterwards follow the partition. */
int cdg_mat_set_partition(cdg_ctx *ctx, cdg_mat *A, const int32_t *part);
int cdg_mat_info(const cdg_mat *A, int32_t *nrows, int32_t *block_n, int32_t *nnzb);

/* y = A x over lblks, chi_blks, bc_blks, hb_blks
 * (chidg_mv, operator_chidg_mv.f90:46-61,117-646). */
int cdg_mv(cdg_ctx *ctx, const cdg_mat *A, const double *x, double *y);
/* Enqueue-only forms for device-resident callers (no reference counterpart):
 * x, y, v, z must be device pointers; the work is queued on the context's
 * stream and the call returns without waiting. */
int cdg_mv_async(cdg_ctx *ctx, const cdg_mat *A, const double *x, double *y);

/* Y_r = A X_r for r < k: k vectors of nrows*n one after another (host or
 * device), each block read once for all k.  Each Y_r has the bits cdg_mv
 * gives for X_r alone.  No reference counterpart beyond chidg_mv per vector. */
int cdg_mv_multi(cdg_ctx *ctx, const cdg_mat *A, int32_t k, const double *X, double *Y);

/* ---- preconditioner: precon_ILU0_t / precon_identity -------------------------
 * create == preconditioner%init (precon_ILU0.f90:64-90)                        */
int cdg_prec_create(cdg_ctx *ctx, const cdg_mat *A, int32_t type, int32_t precision,
                    cdg_prec **out);
/* preconditioner%update(A,b) (precon_ILU0.f90:111-265).  On CDG_ERR_SINGULAR
 * *bad_row (may be NULL) receives the first failing block row. */
int cdg_prec_update(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, int32_t *bad_row);
/* z = preconditioner%apply(A,v) (precon_ILU0.f90:286-411). */
int cdg_prec_apply(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z);
int cdg_prec_apply_async(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z);
/* Z_r = M^-1 V_r for r < k (k vectors one after another, host or device):
 * the row-slice schedule with each block read once for all k (ILU(0));
 * each Z_r has the bits of CDG_SCHEDULE_ROWSLICE's apply of V_r alone.
 * precon_ILU0%apply (precon_ILU0.f90:286-411) per vector in the reference. */
int cdg_prec_apply_multi(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, int32_t k, const double *V, double *Z);
/* preconditioner%tear_down (precon_ILU0.f90:450-470). */
int cdg_prec_destroy(cdg_ctx *ctx, cdg_prec *M);
/* Schedule of the ILU(0) apply (no reference counterpart: the reference's
 * apply is one sequential loop over rows, precon_ILU0.f90:286-411).  Both give
 * the same result for every schedule-independent input order:
 *   CDG_SCHEDULE_LEVELS    one launch per wavefront level and stage
 *   CDG_SCHEDULE_DATAFLOW  one persistent launch, row-to-row dependencies
 *   CDG_SCHEDULE_FOLDED    level schedule with Dinv folded into the upper
 *                          blocks: the first apply after each factor forms
 *                          W = U Dinv (one batched block GEMM over the upper
 *                          blocks, nup more block slots), and the backward
 *                          sweep is one stage per level.  Pays off when one
 *                          factor serves many applies; FP64 and FP32 only.
 *   CDG_SCHEDULE_STREAM    one persistent launch over (row, row slice) tasks:
 *                          the first apply after each factor forms
 *                          W = Dinv U (left fold, one batched block GEMM) and
 *                          copies L and Dinv into a row-slice layout (nlo +
 *                          nup + nrows more block slots), so every task writes
 *                          final values and the backward sweep needs no
 *                          Dinv stage.  All precisions.
 *   CDG_SCHEDULE_ROWSLICE  the tasks of CDG_SCHEDULE_STREAM on the same
 *                          storage, one launch per wavefront level: every task
 *                          reads final parent values written by earlier
 *                          launches (no waits inside a launch).  All precisions.
 *   CDG_SCHEDULE_AUTO      (the default) CDG_SCHEDULE_LEVELS for the first 12
 *                          applies after each factor, CDG_SCHEDULE_ROWSLICE
 *                          from the 13th on, or from the first when the
 *                          previous factor served more than 12 (environment
 *                          CDG_FOLD_AFTER=k sets k; 0 never switches), so long
 *                          solves get the row-slice storage and short ones
 *                          skip its preparation.  */
#define CDG_SCHEDULE_LEVELS   0
#define CDG_SCHEDULE_DATAFLOW 1
#define CDG_SCHEDULE_FOLDED   2
#define CDG_SCHEDULE_AUTO     3
#define CDG_SCHEDULE_STREAM   4
#define CDG_SCHEDULE_ROWSLICE 5
int cdg_prec_set_schedule(cdg_ctx *ctx, cdg_prec *M, int32_t schedule);
/* Launch-shape autotuner (no reference counterpart).  cdg_prec_tune times the
 * apply of M's current factor (level schedules; after cdg_prec_update) under
 * candidate shapes -- column-slice width, threads per workgroup, 16-byte loads
 * in flight -- `reps` calls each, keeps the fastest for M and remembers it for
 * preconditioners of the same block size and precision created later in the
 * process.  cdg_mat_tune does the same for the matvec (threads, loads in
 * flight; remembered per block size).  *best_ms (may be NULL) receives the
 * chosen shape's time per call.  Results do not depend on the shape beyond
 * rounding (the partial sums regroup). */
int cdg_prec_tune(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, int32_t reps, double *best_ms);
int cdg_mat_tune(cdg_ctx *ctx, cdg_mat *A, int32_t reps, double *best_ms);
/* The current shapes: out[4] = {column slices, slice width, threads, loads in
 * flight} for the apply; out[2] = {threads, loads in flight} for the matvec. */
int cdg_prec_launch_shape(const cdg_prec *M, int32_t *out);
int cdg_mat_launch_shape(const cdg_mat *A, int32_t *out);
/* Allocation-free apply.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  z = M^-1 v exactly as
 * cdg_prec_apply_async computes it under CDG_SCHEDULE_ROWSLICE (same kernels,
 * same bits), enqueued on the context's stream; v and z are device pointers.
 * M must be an updated ILU(0) preconditioner whose row-slice storage exists
 * (cdg_prec_set_schedule with CDG_SCHEDULE_ROWSLICE or CDG_SCHEDULE_STREAM
 * allocates it; CDG_ERR_STATE otherwise).  The forward sweep's vector
 * (nrows*n*8 bytes for FP64, nrows*n*4 for FP32/BF16) lives in a library
 * array of CDG_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends;
 * larger systems get CDG_ERR_ARG.  All calls on one GPU share that array:
 * calls that overlap on one GPU (streams of several contexts) race.  The call
 * performs no hipMalloc/hipFree. */
#define CDG_NOALLOC_SCRATCH_BYTES (32 << 20)
int cdg_prec_apply_noalloc(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z);
/* Factor as FP64 for inspection: L (nnzb*n*n, lower kind-0 slots filled,
 * others zero) and Dinv (nrows*n*n), the LD matrix of precon_ILU0_t. */
int cdg_prec_download(cdg_ctx *ctx, const cdg_prec *M, double *L, double *Dinv);

/* ---- linear solver: fgmres_t%solve(A,x,b,M) (type_fgmres.f90:81-310) --------
 * x is overwritten (zero initial guess).  Updates M per params. */
int cdg_fgmres_solve(cdg_ctx *ctx, const cdg_mat *A, cdg_prec *M, const double *b, double *x,
                     const cdg_fgmres_params *params, cdg_fgmres_stats *stats);

/* k right-hand sides of one system with one preconditioner: k runs of
 * fgmres_t%solve (type_fgmres.f90:81-310) on the same A and M, in lockstep.
 * b and x hold k vectors of nrows*n doubles one after another (host or
 * device); stats (may be NULL) receives k entries.  Each vector keeps its own
 * Arnoldi process, rotations, restarts and stopping test (the iterations it
 * takes alone); the steps share one k-vector ILU(0) apply and one k-vector
 * matvec (each block read once for all k) and one host fetch.  M is updated
 * once, per params.  With inner_fgmres the vectors are solved one after
 * another.  No reference counterpart beyond the per-vector solve. */
int cdg_fgmres_solve_multi(cdg_ctx *ctx, const cdg_mat *A, cdg_prec *M, int32_t k, const double *b, double *x,
                           const cdg_fgmres_params *params, cdg_fgmres_stats *stats);

/* Reference defaults of linear_solver_t (type_linear_solver.f90:21-40). */
void cdg_fgmres_default_params(cdg_fgmres_params *params);

/* ---- batch of independent systems --------------------------------------------
 * No reference counterpart (ChiDG solves one system per MPI rank): many
 * systems of uneven sizes and block sizes, nrhs[s] (1..8) right-hand sides
 * each, solved together on one GPU: per right-hand side the outer FGMRES of
 * fgmres_t%solve (type_fgmres.f90:81-310; CGS, inner FGMRES off, nkrylov <=
 * 31), all Arnoldi steps advancing in lockstep.  cdg_batch_create joins the
 * systems of each block size into one block-diagonal matrix (copies of the
 * blocks of A[s], which must be complete), computes its ILU(0) in `precision`
 * (each system's own factor: one factor launch sequence per block size) and
 * allocates the Krylov space of nkrylov steps.  cdg_batch_solve: b[s] and
 * x[s] hold nrhs[s] vectors of system s one after another (host or device);
 * stats (may be NULL) receives one entry per right-hand side, systems in
 * order.  Each Arnoldi step is, per block size, k-vector applies and
 * matvecs over row ranges of systems with like right-hand-side counts (1, 2,
 * 3-4, 5-8 vectors per launch; rows of systems whose lanes have all stopped
 * are skipped), segmented vector kernels over all right-hand sides and one
 * rotation kernel on the device; a cycle of nkrylov steps runs as one
 * captured HIP graph (CDG_BATCH_GRAPH=0: the same launches without the
 * graph; CDG_BATCH_TRACE=1: phase times on stderr).  Each right-hand side
 * gets the iterations of its own solve (the same stopping test, restarts at
 * its own cycle ends).  A solve whose lanes never meet the stopping test runs
 * until nmax (nmax <= 0: without bound, as the reference). */
typedef struct cdg_batch cdg_batch;
int cdg_batch_create(cdg_ctx *ctx, int32_t nsys, cdg_mat *const *A, const int32_t *nrhs, int32_t precision,
                     int32_t nkrylov, cdg_batch **out);
int cdg_batch_solve(cdg_ctx *ctx, cdg_batch *batch, const double *const *b, double *const *x,
                    const cdg_fgmres_params *params, cdg_fgmres_stats *stats);
/* New blocks for the batch's systems (A[s] with the patterns it was created
 * from, e.g. the next Newton step's Jacobians): copied into the joined
 * matrices and refactored, with no allocation.  A = NULL refactors the
 * joined matrices as they are. */
int cdg_batch_update(cdg_ctx *ctx, cdg_batch *batch, cdg_mat *const *A);
int cdg_batch_info(const cdg_batch *batch, int32_t *nlanes, int32_t *ngroups, int64_t *values);
int cdg_batch_destroy(cdg_ctx *ctx, cdg_batch *batch);

/* ---- request loop for parameter-sweep solves ----------------------------------
 * No reference counterpart (ChiDG solves one system per MPI rank); for callers
 * that solve many independent systems on one GPU.  A server runs `nslots`
 * workers, each with its own context and HIP stream.  Requests queue in
 * arrival order; a free worker admits the head of the queue as soon as its
 * device-memory footprint (cdg_server_footprint) fits `hbm_budget` bytes next
 * to the running requests (hbm_budget <= 0: 80% of the free device memory),
 * so work of concurrent requests overlaps on the device.  A request names a
 * synthetic system: uniform hex box nx*ny*nz in ChiDG element order, degree
 * `order` (block n = (order+1)^3 * 5), periodic faces by bit (1: xi, 2: eta,
 * 4: zeta), blocks from cdg_mat_fill_synthetic(seed_A, s_diag, s_off), right-
 * hand side cdg_vec_fill_synthetic(seed_b, 1), ILU(0) in `precision`; the
 * server's FGMRES parameters.  Results do not depend on what runs beside a
 * request.  x_out (may be NULL) receives the solution of the first right-hand
 * side and must stay valid until cdg_server_wait returns for that ticket. */
typedef struct cdg_server cdg_server;
typedef struct {
    int32_t nx, ny, nz, order, periodic, precision;
    uint64_t seed_A, seed_b;
    double s_diag, s_off;
    int32_t nrhs;              /* right-hand sides seed_b, seed_b+1, ... (0 or 1: one); one factor */
} cdg_request;
typedef struct {
    int32_t status;            /* CDG_OK or the failing call's code */
    int32_t niter;             /* FGMRES iterations, summed over the right-hand sides */
    double residual_norm;      /* max over right-hand sides of ||b - A x|| / ||b|| */
    double t_queue;            /* seconds from submit to admission */
    double t_run;              /* seconds from admission to completion */
    double t_latency;          /* seconds from submit to completion */
    double t_factor;           /* ILU(0) update inside t_run */
    int32_t slot;              /* worker that ran the request */
    int32_t factored;          /* ILU(0) updates the request made (a registered
                                  system's resident factor: 0 while its stamp holds) */
} cdg_result;
int cdg_server_create(int32_t device, int32_t nslots, int64_t hbm_budget, const cdg_fgmres_params *params,
                      cdg_server **out);
int cdg_server_submit(cdg_server *s, const cdg_request *req, double *x_out, uint64_t *ticket);
int cdg_server_wait(cdg_server *s, uint64_t ticket, cdg_result *res);
int cdg_server_footprint(const cdg_request *req, const cdg_fgmres_params *params, int64_t *bytes);
int cdg_server_info(cdg_server *s, int32_t *queued, int32_t *active, int32_t *peak_active, int64_t *in_use,
                    int64_t *budget);
const char *cdg_server_last_error(cdg_server *s);
/* Requests on a caller's matrix.  cdg_server_register: A (created and stored
 * by the caller, any context of the server's device) becomes a registered
 * system with an ILU(0) factor of `precision` that stays resident on the
 * device between requests; its bytes are reserved against the budget while
 * registered (CDG_ERR_OOM when they do not fit).  The factor is computed at
 * the first request and again only when A's stamp changed (cdg_mat_upload),
 * the rule of solver_controller_t (type_solver_controller.f90:89-124).  The
 * requests of one system run in submission order on one worker; the caller
 * does not store into A while its requests run, and keeps A alive until
 * cdg_server_unregister returns.  cdg_server_submit_system: nrhs right-hand
 * sides (nrhs * nrows * n doubles, host or device, valid until the wait
 * returns) sharing the factor; x_out (NULL: discarded) receives the nrhs
 * solutions.  cdg_server_unregister waits for the system's queued requests
 * and frees its factor. */
int cdg_server_register(cdg_server *s, const cdg_mat *A, int32_t precision, uint64_t *system);
int cdg_server_submit_system(cdg_server *s, uint64_t system, const double *b, int32_t nrhs, double *x_out,
                             uint64_t *ticket);
int cdg_server_unregister(cdg_server *s, uint64_t system);
int cdg_server_destroy(cdg_server *s);

/* ---- device-resident vectors ---------------------------------------------------
 * chidg_vector_t operations on vectors that stay on the device between calls
 * (type_chidg_vector.f90: the overloaded +, -, * (:1740-2080) and norm
 * (:1237-1300); operator_chidg_dot.f90:30-89 dot), so a Newton loop keeps q,
 * dq and rhs in HBM.  cdg_vec_axpby: y = a x + b y (device pointers);
 * cdg_vec_dot / cdg_vec_norm: host or device vectors, result on the host. */
/* cdg_vec_create: a zeroed device vector of len doubles from the context's
 * memory (cdg_vec_destroy returns it); cdg_vec_copy: len doubles between any
 * host / device pointers (upload, download, device copy), synchronous. */
int cdg_vec_create(cdg_ctx *ctx, int64_t len, double **out);
int cdg_vec_destroy(cdg_ctx *ctx, double *v);
int cdg_vec_copy(cdg_ctx *ctx, int64_t len, const double *src, double *dst);
int cdg_vec_axpby(cdg_ctx *ctx, int64_t len, double a, const double *x, double b, double *y);
int cdg_vec_dot(cdg_ctx *ctx, int64_t len, const double *x, const double *y, double *out);
int cdg_vec_norm(cdg_ctx *ctx, int64_t len, const double *x, double *out);

/* ---- DG Euler first-Newton-step system (host code) --------------------------
 * The system ChiDG assembles for the first Newton step of a steady Euler case
 * (mod_spatial.f90:40-197 residual and Jacobian, type_newton.f90:320-351
 * pseudo-transient term) on the periodic uniform hex box [0,1]^3 with
 * nx = ny = nz >= 2 elements per direction and order 0..3: 5 conservative
 * fields, tensor Legendre modal basis (n = 5 (order+1)^3, block index
 * field*nterms + mode), Gauss-Legendre quadrature with order+2 points,
 * central flux plus Roe dissipation with the entropy fix of
 * euler_roe_operator.f90:200-215, mass/dtau on the diagonal blocks with
 * dtau = cfl h / (mean |u| + mean c) (type_fluid_pseudo_timestep.f90:95-155).
 * Blocks follow the hex-box pattern (all directions periodic; the wrap
 * neighbours are coupling blocks).  Vectors are nrows*n, element-major.
 * nthreads <= 0: the host's cores (at most 16).  Cubes only: nx = ny = nz in
 * 2..256, order 0..3, else CDG_ERR_ARG; CDG_ERR_NAN when the state gives a
 * non-finite residual or Jacobian (a negative pressure leaves the pseudo-time
 * step undefined); CDG_ERR_OOM when a host allocation fails. */
/* q: freestream rho = 1, p = 1/gamma, u = (mach, 0, 0) on mode 0, plus
 * perturb * N(0,1) (seeded) on every coefficient. */
int cdg_euler_state(int32_t nx, int32_t ny, int32_t nz, int32_t order, double mach, uint64_t seed, double perturb,
                    double *q);
/* R(q): sum_faces int phi Fhat.n - int grad(phi).F over each element. */
int cdg_euler_residual(int32_t nx, int32_t ny, int32_t nz, int32_t order, const double *q, double *R,
                       int32_t nthreads);
/* R(q) on the context's GPU: the same discretization, one workgroup per
 * element (the spatial residual assembly of mod_spatial.f90:40-197); q and R
 * host or device; equal to cdg_euler_residual to rounding (sums in another
 * order).  CDG_ERR_NAN for a non-finite residual. */
int cdg_euler_residual_gpu(cdg_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, int32_t order, const double *q, double *R);
/* blocks (nnzb*n*n, column-major per block) = dR/dq + mass/dtau (cfl > 0);
 * R (may be NULL) receives R(q). */
int cdg_euler_jacobian(int32_t nx, int32_t ny, int32_t nz, int32_t order, const double *q, double cfl,
                       double *blocks, double *R, int32_t nthreads);

#ifdef __cplusplus
}
#endif

#endif
