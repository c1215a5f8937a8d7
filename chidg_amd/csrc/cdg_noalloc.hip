// Allocation-free ILU(0) apply (no counterpart in the reference): for callers
// that cannot allocate device memory -- code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc.
//
// cdg_prec_apply_noalloc runs the row-slice level schedule (the tasks of
// CDG_SCHEDULE_ROWSLICE, bit for bit the same kernels and sums) with its one
// scratch buffer -- the forward sweep's vector y, nrows * n values -- in the
// file-scope device array `scratch` below instead of memory of the
// preconditioner.  Each GPU holds its own copy of the array until the process
// ends; its device address is looked up once per device (noalloc_init, called
// by cdg_ctx_create with the device current) and only that address is passed
// to kernels.  Every call on one GPU shares the array: each call writes every
// y value before reading it, so nothing needs clearing, but calls that overlap
// on one GPU (streams of several contexts) race on it.  The call allocates
// nothing: the preconditioner's row-slice storage must already exist
// (cdg_prec_set_schedule with CDG_SCHEDULE_ROWSLICE or CDG_SCHEDULE_STREAM
// allocates it), and the first call after each update forms W = Dinv U and the
// row-slice copies in that storage (kernels only).
#include "cdg_internal.h"

#include <mutex>

static __device__ __align__(256) unsigned char scratch[CDG_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};        // device address of `scratch`, per device
hipError_t g_status[kMaxDevices] = {};    // the lookup's result, per device
bool g_done[kMaxDevices] = {};
}  // namespace

// Once per device, with `device` current.
int noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return CDG_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        void *p = nullptr;
        g_status[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[device] = g_status[device] == hipSuccess ? p : nullptr;
        g_done[device] = true;
    }
    return g_status[device] == hipSuccess ? CDG_OK : CDG_ERR_HIP;
}

extern "C" int cdg_prec_apply_noalloc(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    if (!ctx || !M || !A || !v || !z) return CDG_ERR_ARG;
    if (!is_device_ptr(v) || !is_device_ptr(z))
        return cdg_fail(ctx, CDG_ERR_ARG, "cdg_prec_apply_noalloc needs device pointers");
    if (M->type != CDG_PREC_ILU0) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "cdg_prec_apply_noalloc: ILU(0) only");
    if (A->nrows != M->nrows || A->n != M->n) return cdg_fail(ctx, CDG_ERR_ARG, "matrix does not match preconditioner");
    const size_t need = (size_t)M->nrows * M->n * (M->precision == CDG_FP64 ? 8 : 4);
    if (need > (size_t)CDG_NOALLOC_SCRATCH_BYTES)
        return cdg_fail(ctx, CDG_ERR_ARG,
                        "cdg_prec_apply_noalloc: the forward vector needs %zu bytes, the scratch array holds %d",
                        need, CDG_NOALLOC_SCRATCH_BYTES);
    if (!M->factored) return cdg_fail(ctx, CDG_ERR_STATE, "ILU0 apply before update");
    if (!M->rs_tasks)
        return cdg_fail(ctx, CDG_ERR_STATE,
                        "cdg_prec_apply_noalloc: no row-slice storage (set CDG_SCHEDULE_ROWSLICE first)");
    int e = noalloc_init(ctx->device);
    if (e) return cdg_fail(ctx, e, "scratch array address: %s", hipGetErrorString(g_status[ctx->device]));
    void *y = g_scratch[ctx->device];
    if (!M->rs_ready) {   // W = Dinv U and the row-slice copies of this factor (kernels, no allocation)
        e = launch_fold_rs(ctx, M, A);
        if (!e) e = rs_relayout(ctx, M);
        if (e) return e;
        M->rs_ready = 1;
    }
    return launch_ilu_apply_rsl_buf(ctx, M, v, z, y);
}
