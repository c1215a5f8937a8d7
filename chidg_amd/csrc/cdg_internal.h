// Internal types of the gfx950 linear-solve library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/chidg_hip.h"

#define CDG_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return cdg_fail(ctx, CDG_ERR_HIP, "%s: %s (%s:%d)", #expr, \
                                              hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

int cdg_fail(const cdg_ctx *ctx, int code, const char *fmt, ...);

// 16-bit brain float storage
typedef uint16_t cdg_bf16;

// Device memory owned by a context (freed in the owner's destructor path).
struct DevMem {
    void *p = nullptr;
    size_t bytes = 0;
};

struct cdg_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    mutable std::string err;
    int64_t mem_used = 0;              // bytes held by live objects
    int64_t mem_cached = 0;            // bytes freed by objects, kept for reuse
    std::unordered_map<void *, size_t> alloc_size;
    std::multimap<size_t, void *> cache;
    // reduction scratch (device) and pinned host mirror
    double *d_part = nullptr;     // partial sums
    size_t part_cap = 0;          // doubles
    double *d_small = nullptr;    // small results (dots, norms)
    double *h_small = nullptr;    // pinned host copy
    size_t small_cap = 0;
    // staging for host pointers passed to the ABI
    double *d_stage[3] = {nullptr, nullptr, nullptr};
    size_t stage_cap[3] = {0, 0, 0};
    // completion event of ctx_wait (the solver's per-step waits poll it)
    hipEvent_t done_ev = nullptr;
    hipEvent_t order_ev = nullptr;     // cdg_ctx_wait_stream / cdg_ctx_signal_stream
    bool spin = true;
    bool counted = false;              // included in the live-context count (stream sweep grid)
};

// Waits until the context's stream has finished its work: polls a completion
// event (a few microseconds of latency instead of a blocking wait's tens), and
// falls back to a blocking wait after 5 ms.  CDG_SPIN=0 always blocks.
int ctx_wait(cdg_ctx *ctx);

// Block-row streaming launch shape for one block size / element width.
struct StreamShape {
    int T = 0;     // threads per workgroup (V*T is a multiple of n)
    int dc = 0;    // column advance per iteration = V*T/n
};
// Launch shape of the fixed-row walk; T <= tmax is a multiple of n/gcd(n,V)
// chosen for lane utilization, preferring sizes near `target`.
StreamShape stream_shape(int n, int V, int target, int tmax = 1024);

struct cdg_mat {
    int32_t nrows = 0, n = 0, nnzb = 0;
    int64_t ld = 0;                    // elements per block slot (padded, multiple of 64)
    double *blocks = nullptr;          // nnzb * ld, device
    int32_t *d_row_ptr = nullptr, *d_col = nullptr;
    std::vector<int32_t> row_ptr, col, kind;
    std::vector<int32_t> part;         // rank of each row (empty: one rank), see cdg_mat_set_partition
    int32_t max_row_blocks = 0;
    uint64_t stamp = 0;
    // kind-0 structure (type_domain_matrix.f90:295-302, 592-627)
    std::vector<int32_t> diag, lo_ptr, lo_blk, lo_col, lo_tr, up_ptr, up_blk, up_col, up_tr;
    int32_t struct_err = CDG_OK;       // CDG_ERR_NO_TRANSPOSE if the kind-0 pattern is not symmetric
    std::vector<int32_t> flevel, blevel;
    int32_t nflev = 0, nblev = 0;
    StreamShape shape64;
    int32_t mv_unroll = 8;             // 16-byte loads in flight per thread (matvec)
};

struct cdg_prec {
    int32_t type = CDG_PREC_IDENTITY;
    int32_t jacobi = 0;                // created as CDG_PREC_JACOBI (ILU(0) with no off-diagonal blocks)
    int32_t precision = CDG_FP64;
    int32_t nrows = 0, n = 0, nlo = 0, nup = 0;
    int64_t ld = 0;                    // elements per factor slot
    // factor storage (element type by precision)
    void *L = nullptr;                 // nlo slots
    void *U = nullptr;                 // nup slots (FP32/BF16 copies of A's upper blocks)
    // nup slots, FP64/FP32: W_u = A_{up_blk[u]} Dinv_{up_col[u]}, the upper
    // blocks with the Dinv of their column folded in (the level schedule's
    // backward sweep is then one stage per level); null for BF16
    void *W = nullptr;
    int32_t w_ready = 0;               // W holds the current factor's products
    // CDG_APPLY_AUTO: W is formed once the factor has served fold_after
    // applies (0: never), or at its first apply when the previous factor
    // served more than fold_after (consecutive solves of a Newton loop take
    // similar iteration counts), so long solves get the shorter backward sweep
    int32_t fold_after = 12, napply = 0, prev_napply = 0;
    int32_t *d_up_all = nullptr;       // 0..nup-1 (the W jobs)
    void *Dinv = nullptr;              // nrows slots
    float *L32 = nullptr, *D32 = nullptr;   // FP32 work for the BF16 variant
    void *inv_w = nullptr;             // working slots of the global-memory block inverse
    int32_t *d_redo = nullptr;         // per row: block inverse must be redone with partial pivoting
    // device index arrays
    int32_t *d_lo_ptr = nullptr, *d_lo_blk = nullptr, *d_lo_col = nullptr, *d_lo_tr = nullptr;
    int32_t *d_lo_tru = nullptr;       // U slot of lo_tr[q] (FP32/BF16: the upper blocks' copies)
    int32_t *d_up_ptr = nullptr, *d_up_idx = nullptr, *d_up_col = nullptr, *d_up_blk = nullptr;
    int32_t *d_diag = nullptr;
    int32_t *d_err = nullptr;          // first singular row (factor)
    double *d_part = nullptr;          // apply partial sums: 3 arrays of part_len doubles
    int64_t part_len = 0;              // nrows * slices * n
    int64_t part_cap = 0;              // doubles allocated at d_part
    // forward levels: rows grouped by level (factor and forward sweep); lower
    // blocks of those rows grouped the same way (factor phase 1)
    std::vector<int32_t> lev_lo_ptr, lev_row_ptr;
    int32_t *d_lev_lo = nullptr, *d_lev_rows = nullptr;
    int32_t nlev = 0;
    // backward levels (backward sweep)
    std::vector<int32_t> blev_row_ptr;
    int32_t *d_blev_rows = nullptr;
    int32_t nblev = 0;
    std::vector<int32_t> h_lev_rows, h_blev_rows;   // host copies of the level orders
    std::vector<int32_t> h_lo_ptr, h_lo_col, h_up_ptr, h_up_col;   // host copies of the kind-0 lists
    int32_t maxdep = 1;               // max lower/upper blocks of a row
    // apply launch shape: each block's columns are cut into `slices` ranges of
    // `slice_w` columns; shape[precision] is the fixed-row walk for that element
    int32_t slices = 1, slice_w = 1, unroll = 4;
    StreamShape shape[3];
    // apply schedule: CDG_APPLY_LEVELS (one launch per level and stage),
    // CDG_APPLY_DATAFLOW (one persistent launch, k_dfapply.hip) or
    // CDG_APPLY_FOLDED (levels, backward over W = U Dinv) or CDG_APPLY_AUTO
    // (levels, then folded once the factor has served fold_after applies)
    int32_t apply_mode = 0;
    int32_t df_T = 0, df_dc = 0, df_rounds = 0, df_w = 0, df_S = 0, df_grid = 0, df_ntasks = 0;
    size_t df_lds = 0;
    int32_t *d_tasks = nullptr;        // task codes in ticket order
    int32_t *d_cnt = nullptr;          // [3][nrows] finished slices (cumulative)
    unsigned long long *d_ticket = nullptr;
    int32_t *d_dferr = nullptr;        // raised by a sweep that gave up waiting
    int32_t df_epoch = 0;              // calls since the counters were zeroed
    int32_t df_spin = 1 << 21;         // polls before a dependency wait gives up
    // stream schedule (k_rsapply.hip): the factor in the row-slice layout
    // (slices of rs_w rows, slot pitch rs_slot elements), W = Dinv U, the task
    // list and the tagged exchange buffers
    void *Lrs = nullptr, *Wrs = nullptr, *Drs = nullptr;
    int32_t *rs_tasks = nullptr;       // device task records
    int32_t *d_up_row = nullptr;       // row of each upper entry
    void *rs_yb = nullptr, *rs_zb = nullptr;
    void *rs_y = nullptr;              // forward result of the row-slice level schedule
    void *ym = nullptr;                // forward results of k-vector applies (cdg_prec_apply_multi)
    size_t ym_bytes = 0;
    uint32_t *d_rs_ctl = nullptr;      // error word + ticket counters of the sweep
    uint32_t rs_epoch = 0;             // tag of the last call
    int32_t rs_ready = 0;              // Lrs/Wrs/Drs hold the current factor
    int32_t rs_ok = 1;                 // 0: the pattern does not fit the stream schedule (AUTO keeps levels)
    int32_t rs_w = 0, rs_S = 0, rs_T = 0, rs_dc = 0, rs_R = 0, rs_EPT = 0, rs_ntasks = 0, rs_per_cu = 0;
    int32_t rs_spin = 1 << 20;
    int32_t rs_classes = 64;           // ticket classes of the sweep
    int64_t rs_SL = 0, rs_slot = 0;
    size_t rs_lds = 0;
    uint64_t stamp = 0;
    int32_t factored = 0;
    std::vector<DevMem> owned;
};

// ---- kernel launchers (defined in the .hip files) ----------------------------
int launch_mv(cdg_ctx *ctx, const cdg_mat *A, const double *x, double *y);
int launch_ilu_apply(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z);
int launch_ilu_apply_df(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z);
int df_configure(cdg_ctx *ctx, cdg_prec *M);
int df_reset(cdg_ctx *ctx, cdg_prec *M);
enum { CDG_APPLY_LEVELS = 0, CDG_APPLY_DATAFLOW = 1, CDG_APPLY_FOLDED = 2, CDG_APPLY_AUTO = 3, CDG_APPLY_STREAM = 4,
       CDG_APPLY_ROWSLICE = 5 };
// stream schedule (k_rsapply.hip; the fold in k_factor.hip)
int rs_setup(cdg_ctx *ctx, cdg_prec *M);
int rs_relayout(cdg_ctx *ctx, cdg_prec *M);
int launch_fold_rs(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A);
int launch_ilu_apply_rs(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z);
int launch_ilu_apply_rs_buf(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z, uint64_t *yb, uint64_t *zb,
                            uint32_t *ctl, uint32_t tag);
int rs_ctl_words();
int launch_ilu_apply_rsl(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z);
int launch_ilu_apply_rsl_buf(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z, void *y);
int noalloc_init(int device);
// W = Dinv U and the row-slice copies of M's current factor (once per factor)
int rs_prepare(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A);
int launch_ilu_apply_rs_multi(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z, void *y, int64_t stride, int k,
                              const int32_t *row_k = nullptr, int row_lo = 0, int row_hi = -1);
int launch_mv_multi(cdg_ctx *ctx, const cdg_mat *A, const double *X, double *Y, int64_t stride, int k,
                    const int32_t *row_k = nullptr, int row_lo = 0, int row_hi = -1);
bool is_device_ptr(const void *p);
int rs_slice_rows(int n, int wmax);
void rs_ctx_count(int delta);
// W_u = A_{up_blk[u]} Dinv_{up_col[u]} for all upper blocks (after a factor)
int launch_fold_all(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A);
// Uniform hex-box pattern in ChiDG row order (cdg_serve.hip); periodic: bit d
// for direction d; face (optional) 1..6 face id per entry, 7 the diagonal
void cdg_hex_box(int nx, int ny, int nz, int periodic, std::vector<int32_t> &row_ptr, std::vector<int32_t> &col,
                 std::vector<int32_t> &kind, std::vector<int32_t> *face = nullptr);
int launch_ilu_factor(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, int32_t *bad_row);
int launch_fill_synthetic(cdg_ctx *ctx, cdg_mat *A, uint64_t seed, double s_diag, double s_off);
int launch_vec_synthetic(cdg_ctx *ctx, int64_t len, uint64_t seed, double scale, double *v);
int launch_widen(cdg_ctx *ctx, const void *src, int precision, double *dst, int64_t count);

// Krylov vector kernels (k_krylov.hip); all on ctx->stream, device pointers
// d_gate (may be null): a device flag; 0.0 skips the kernels (a gated pass)
int kv_multidot(cdg_ctx *ctx, int64_t len, const double *w, const double *V, int64_t vstride, int j,
                double *d_out /* j doubles */, const double *d_gate = nullptr);
int kv_sumsq(cdg_ctx *ctx, int64_t len, const double *w, double *d_out);
// w = w - sum_i h_i V_i (i ascending), h on device; also d_nrm2 = ||w_new||^2 when non-null.
// d_flag (may be null) receives kv_reorth_gate(d_h, j) computed in the same launches.
int kv_cgs_update(cdg_ctx *ctx, int64_t len, double *w, const double *V, int64_t vstride, int j,
                  const double *d_h, double *d_nrm2, const double *d_gate = nullptr, double *d_flag = nullptr);
// *d_flag = 1.0 when sum_{i<jj} |d_h[i]| / sqrt(d_h[jj]) >= 1 (second CGS pass), else 0.0
int kv_reorth_gate(cdg_ctx *ctx, const double *d_h, int jj, double *d_flag);
int kv_axpy_dev(cdg_ctx *ctx, int64_t len, double *w, const double *v, const double *d_h);
int kv_div(cdg_ctx *ctx, int64_t len, const double *w, double h, double *out);
int kv_div_inplace(cdg_ctx *ctx, int64_t len, double *w, double h);
int kv_sub(cdg_ctx *ctx, int64_t len, const double *b, const double *Ax, double *r);   // r = b - Ax
int kv_add(cdg_ctx *ctx, int64_t len, double *z, const double *dz);                   // z += dz
int kv_combine(cdg_ctx *ctx, int64_t len, double *x, const double *x0, const double *Z, int64_t zstride,
               int k, const double *h_y /* host */);                                    // x = x0 + sum y_i Z_i
int kv_zero(cdg_ctx *ctx, int64_t len, double *x);
int kv_axpby(cdg_ctx *ctx, int64_t len, double a, const double *x, double b, double *y);   // y = a x + b y
int launch_euler_residual(cdg_ctx *ctx, int nx, int ny, int nz, int order, const double *q, double *R);
int kv_copy(cdg_ctx *ctx, int64_t len, const double *src, double *dst);

int ctx_ensure_small(cdg_ctx *ctx, size_t n);
int ctx_norm2(cdg_ctx *ctx, int64_t len, const double *v, double *out);   // ||v|| (synchronizes)
int64_t slot_elems(int n);   // elements per block slot: n*n rounded up to 64
int ctx_ensure_part(cdg_ctx *ctx, size_t n);
int dev_alloc(cdg_ctx *ctx, void **p, size_t bytes);
void dev_free(cdg_ctx *ctx, void *p, size_t bytes);
