// C ABI of the gfx950 linear-solve library: contexts, block-sparse matrices,
// ILU(0)/identity preconditioners and restarted FGMRES.
#include "cdg_internal.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>

// ---------------------------------------------------------------------------
// errors and memory
// ---------------------------------------------------------------------------
int cdg_fail(const cdg_ctx *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

// Device memory of a context.  Freed blocks stay cached in the context and
// are handed out again to later requests of up to their size (at most twice
// the request): every use of a context's memory is ordered on its one
// stream, and hipFree would synchronize the whole device, stalling the other
// contexts' work.  The cache is released when the context is destroyed, or
// when an allocation fails.
static void ctx_flush_cache(cdg_ctx *ctx) {
    if (ctx->cache.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->cache) (void)hipFree(kv.second);
    ctx->cache.clear();
    ctx->mem_cached = 0;
}

int dev_alloc(cdg_ctx *ctx, void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return CDG_OK;
    bytes = (bytes + 255) / 256 * 256;
    auto it = ctx->cache.lower_bound(bytes);
    if (it != ctx->cache.end() && it->first <= 2 * bytes) {
        *p = it->second;
        ctx->alloc_size[*p] = it->first;
        ctx->mem_cached -= (int64_t)it->first;
        ctx->mem_used += (int64_t)it->first;
        ctx->cache.erase(it);
        return CDG_OK;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx_flush_cache(ctx);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return cdg_fail(ctx, CDG_ERR_OOM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    ctx->alloc_size[*p] = bytes;
    ctx->mem_used += (int64_t)bytes;
    return CDG_OK;
}

void dev_free(cdg_ctx *ctx, void *p, size_t) {
    if (!p) return;
    auto it = ctx->alloc_size.find(p);
    if (it == ctx->alloc_size.end()) return;
    const size_t bytes = it->second;
    ctx->alloc_size.erase(it);
    ctx->mem_used -= (int64_t)bytes;
    ctx->mem_cached += (int64_t)bytes;
    ctx->cache.emplace(bytes, p);
}

int ctx_norm2(cdg_ctx *ctx, int64_t len, const double *v, double *out) {
    int e = kv_sumsq(ctx, len, v, ctx->d_small);
    if (!e) {
        CDG_HIP_TRY(hipMemcpyAsync(ctx->h_small, ctx->d_small, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
        *out = std::sqrt(ctx->h_small[0]);
    }
    return e;
}

int ctx_ensure_part(cdg_ctx *ctx, size_t n) {
    if (n <= ctx->part_cap) return CDG_OK;
    CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    dev_free(ctx, ctx->d_part, ctx->part_cap * sizeof(double));
    ctx->d_part = nullptr;
    ctx->part_cap = 0;
    size_t cap = std::max<size_t>(n, 4096);
    int e = dev_alloc(ctx, (void **)&ctx->d_part, cap * sizeof(double));
    if (e) return e;
    ctx->part_cap = cap;
    return CDG_OK;
}

int ctx_ensure_small(cdg_ctx *ctx, size_t n) {
    if (n <= ctx->small_cap) return CDG_OK;
    CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    dev_free(ctx, ctx->d_small, ctx->small_cap * sizeof(double));
    if (ctx->h_small) (void)hipHostFree(ctx->h_small);
    ctx->d_small = nullptr;
    ctx->h_small = nullptr;
    ctx->small_cap = 0;
    size_t cap = std::max<size_t>(n, 256);
    int e = dev_alloc(ctx, (void **)&ctx->d_small, cap * sizeof(double));
    if (e) return e;
    if (hipHostMalloc((void **)&ctx->h_small, cap * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return cdg_fail(ctx, CDG_ERR_OOM, "hipHostMalloc failed");
    ctx->small_cap = cap;
    return CDG_OK;
}

bool is_device_ptr(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice;
}

namespace {

std::atomic<uint64_t> g_stamp{0};

// Device view of a vector argument: device pointers pass through, host
// pointers go through a per-slot staging buffer.
struct VecArg {
    double *dev = nullptr;
    double *host = nullptr;
    int64_t len = 0;
};

int stage_in(cdg_ctx *ctx, int slot, const double *p, int64_t len, VecArg &a) {
    a.len = len;
    if (is_device_ptr(p)) { a.dev = const_cast<double *>(p); return CDG_OK; }
    a.host = const_cast<double *>(p);
    size_t need = (size_t)std::max<int64_t>(len, 1);
    if (ctx->stage_cap[slot] < need) {
        CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
        dev_free(ctx, ctx->d_stage[slot], ctx->stage_cap[slot] * sizeof(double));
        ctx->d_stage[slot] = nullptr;
        ctx->stage_cap[slot] = 0;
        int e = dev_alloc(ctx, (void **)&ctx->d_stage[slot], need * sizeof(double));
        if (e) return e;
        ctx->stage_cap[slot] = need;
    }
    a.dev = ctx->d_stage[slot];
    return CDG_OK;
}

int upload_in(cdg_ctx *ctx, const VecArg &a) {
    if (a.host && a.len) CDG_HIP_TRY(hipMemcpyAsync(a.dev, a.host, sizeof(double) * a.len, hipMemcpyHostToDevice, ctx->stream));
    return CDG_OK;
}

int finish_out(cdg_ctx *ctx, const VecArg &a) {
    if (a.host && a.len) CDG_HIP_TRY(hipMemcpyAsync(a.host, a.dev, sizeof(double) * a.len, hipMemcpyDeviceToHost, ctx->stream));
    CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return CDG_OK;
}

}  // namespace

int64_t slot_elems(int n) { return ((int64_t)n * n + 63) / 64 * 64; }

// ---- launch shapes and their autotuner ----------------------------------------
// The apply's shape: column-slice width, threads per workgroup (target of
// stream_shape) and 16-byte loads in flight; the matvec's: threads and loads
// in flight.  Defaults are the measured best at n = 135; cdg_prec_tune /
// cdg_mat_tune time candidates on a live factor / matrix and keep the fastest,
// and remember it per (block size, precision) for objects created afterwards.
namespace {
struct ApplyShape {
    int sw, target, unroll;
};
struct MvShape {
    int target, unroll;
};
std::mutex g_tune_mu;
std::map<std::pair<int, int>, ApplyShape> g_apply_tuned;   // (n, precision)
std::map<int, MvShape> g_mv_tuned;                          // n

// measured at n = 135: 32-column slices; 8 loads in flight for FP64, 4 for FP32/BF16
ApplyShape apply_default(int n, int precision) {
    const bool p2 = n > 64 && n <= 192;
    return {p2 ? 32 : 40, 256, p2 && precision == CDG_FP64 ? 8 : 4};
}
MvShape mv_default() { return {540, 8}; }

bool apply_tuned(int n, int precision, ApplyShape *out) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_apply_tuned.find({n, precision});
    if (it == g_apply_tuned.end()) return false;
    *out = it->second;
    return true;
}

bool mv_tuned(int n, MvShape *out) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_mv_tuned.find(n);
    if (it == g_mv_tuned.end()) return false;
    *out = it->second;
    return true;
}

// Column slices are multiples of 8 columns, so every slice starts 16-byte
// aligned for all element widths.  False if the shape does not fit (LDS,
// threads); M is then unchanged.
bool apply_shape_set(cdg_prec *M, int precision, const ApplyShape &a) {
    const int n = M->n;
    int slices = std::max(1, (n + a.sw - 1) / a.sw);
    const int slice_w = ((n + slices - 1) / slices + 7) / 8 * 8;
    slices = (n + slice_w - 1) / slice_w;
    const StreamShape sh[3] = {stream_shape(n, 2, a.target), stream_shape(n, 4, a.target), stream_shape(n, 8, a.target)};
    const int V = precision == CDG_FP64 ? 2 : (precision == CDG_FP32 ? 4 : 8);
    const size_t accsz = precision == CDG_FP64 ? 8 : 4;
    const size_t lds = accsz * ((size_t)M->maxdep * slice_w + (size_t)V * sh[precision].T);
    if (sh[precision].T == 0 || lds > 64 * 1024) return false;
    M->slices = slices;
    M->slice_w = slice_w;
    M->unroll = a.unroll == 8 ? 8 : 4;
    for (int i = 0; i < 3; ++i) M->shape[i] = sh[i];
    M->part_len = (int64_t)M->nrows * slices * n;
    return true;
}
}  // namespace

namespace {

// Allocates and fills a device array.  Waits for the copy: callers pass
// temporaries, and an asynchronous copy from pageable memory may read its
// source after the call returns.
template <typename T>
int upload_array(cdg_ctx *ctx, T **d, const std::vector<T> &h, std::vector<DevMem> *owned) {
    size_t bytes = std::max<size_t>(1, h.size()) * sizeof(T);
    int e = dev_alloc(ctx, (void **)d, bytes);
    if (e) return e;
    if (owned) owned->push_back({*d, bytes});
    if (!h.empty()) {
        CDG_HIP_TRY(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
        CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return CDG_OK;
}

// kind-0 structure, transposes and wavefront levels of a pattern
void analyze(cdg_mat *A) {
    const int nr = A->nrows;
    A->diag.assign(nr, -1);
    A->lo_ptr.assign(nr + 1, 0);
    A->up_ptr.assign(nr + 1, 0);
    A->lo_blk.clear(); A->lo_col.clear(); A->up_blk.clear(); A->up_col.clear();
    A->struct_err = CDG_OK;
    for (int i = 0; i < nr; ++i) {
        for (int b = A->row_ptr[i]; b < A->row_ptr[i + 1]; ++b) {
            if (A->kind[b] != CDG_KIND_LOCAL) continue;
            const int j = A->col[b];
            if (j == i) { if (A->diag[i] < 0) A->diag[i] = b; }
            else if (!A->part.empty() && A->part[j] != A->part[i]) continue;   // parent on another rank
            else if (j < i) { A->lo_blk.push_back(b); A->lo_col.push_back(j); }
            else { A->up_blk.push_back(b); A->up_col.push_back(j); }
        }
        A->lo_ptr[i + 1] = (int32_t)A->lo_blk.size();
        A->up_ptr[i + 1] = (int32_t)A->up_blk.size();
        if (A->diag[i] < 0) A->struct_err = CDG_ERR_ARG;
    }
    auto find_t = [&](int row, int parent) {
        for (int b = A->row_ptr[row]; b < A->row_ptr[row + 1]; ++b)
            if (A->kind[b] == CDG_KIND_LOCAL && A->col[b] == parent) return b;
        return -1;
    };
    A->lo_tr.resize(A->lo_blk.size());
    A->up_tr.resize(A->up_blk.size());
    for (int i = 0; i < nr; ++i) {
        for (int q = A->lo_ptr[i]; q < A->lo_ptr[i + 1]; ++q) {
            A->lo_tr[q] = find_t(A->lo_col[q], i);
            if (A->lo_tr[q] < 0 && A->struct_err == CDG_OK) A->struct_err = CDG_ERR_NO_TRANSPOSE;
        }
        for (int u = A->up_ptr[i]; u < A->up_ptr[i + 1]; ++u) {
            A->up_tr[u] = find_t(A->up_col[u], i);
            if (A->up_tr[u] < 0 && A->struct_err == CDG_OK) A->struct_err = CDG_ERR_NO_TRANSPOSE;
        }
    }
    A->flevel.assign(nr, 0);
    A->blevel.assign(nr, 0);
    A->nflev = nr ? 1 : 0;
    A->nblev = nr ? 1 : 0;
    for (int i = 0; i < nr; ++i) {
        int l = 0;
        for (int q = A->lo_ptr[i]; q < A->lo_ptr[i + 1]; ++q) l = std::max(l, A->flevel[A->lo_col[q]] + 1);
        A->flevel[i] = l;
        A->nflev = std::max(A->nflev, l + 1);
    }
    for (int i = nr - 1; i >= 0; --i) {
        int l = 0;
        for (int u = A->up_ptr[i]; u < A->up_ptr[i + 1]; ++u) l = std::max(l, A->blevel[A->up_col[u]] + 1);
        A->blevel[i] = l;
        A->nblev = std::max(A->nblev, l + 1);
    }
    A->max_row_blocks = 0;
    for (int i = 0; i < nr; ++i) A->max_row_blocks = std::max(A->max_row_blocks, A->row_ptr[i + 1] - A->row_ptr[i]);
}

std::vector<int32_t> order_by_level(const std::vector<int32_t> &lev, int nlev, std::vector<int32_t> *ptr_out) {
    std::vector<int32_t> ptr(nlev + 1, 0), rows(lev.size());
    for (int32_t l : lev) ptr[l + 1]++;
    for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (size_t i = 0; i < lev.size(); ++i) rows[fill[lev[i]]++] = (int32_t)i;
    if (ptr_out) *ptr_out = ptr;
    return rows;
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
int ctx_wait(cdg_ctx *ctx) {
    if (!ctx->spin || !ctx->done_ev) {
        CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
        return CDG_OK;
    }
    CDG_HIP_TRY(hipEventRecord(ctx->done_ev, ctx->stream));
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t q;
    while ((q = hipEventQuery(ctx->done_ev)) == hipErrorNotReady) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(5)) {
            q = hipEventSynchronize(ctx->done_ev);
            break;
        }
    }
    if (q != hipSuccess) return cdg_fail(ctx, CDG_ERR_HIP, "%s", hipGetErrorString(q));
    return CDG_OK;
}

extern "C" int cdg_ctx_create(int32_t device, cdg_ctx **out) {
    if (!out) return CDG_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { (void)hipGetLastError(); return CDG_ERR_HIP; }
    if (device < 0 || device >= ndev) return CDG_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return CDG_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CDG_ERR_HIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return CDG_ERR_UNSUPPORTED;
    cdg_ctx *ctx = new cdg_ctx();
    ctx->device = device;
    ctx->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return CDG_ERR_HIP; }
    if (hipEventCreateWithFlags(&ctx->done_ev, hipEventDisableTiming) != hipSuccess) {
        ctx->done_ev = nullptr;
        cdg_ctx_destroy(ctx);
        return CDG_ERR_HIP;
    }
    if (const char *sp = std::getenv("CDG_SPIN")) ctx->spin = std::atoi(sp) != 0;
    if (ctx_ensure_small(ctx, 4096) || ctx_ensure_part(ctx, 1 << 16)) { cdg_ctx_destroy(ctx); return CDG_ERR_OOM; }
    if (noalloc_init(device) != CDG_OK) { cdg_ctx_destroy(ctx); return CDG_ERR_HIP; }   // scratch array address
    rs_ctx_count(+1);
    ctx->counted = true;
    *out = ctx;
    return CDG_OK;
}

extern "C" int cdg_ctx_destroy(cdg_ctx *ctx) {
    if (!ctx) return CDG_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    dev_free(ctx, ctx->d_part, ctx->part_cap * sizeof(double));
    dev_free(ctx, ctx->d_small, ctx->small_cap * sizeof(double));
    if (ctx->h_small) (void)hipHostFree(ctx->h_small);
    for (int s = 0; s < 3; ++s) dev_free(ctx, ctx->d_stage[s], ctx->stage_cap[s] * sizeof(double));
    for (auto &kv : ctx->alloc_size) (void)hipFree(kv.first);   // blocks of objects never destroyed
    ctx->alloc_size.clear();
    ctx_flush_cache(ctx);
    if (ctx->done_ev) (void)hipEventDestroy(ctx->done_ev);
    if (ctx->order_ev) (void)hipEventDestroy(ctx->order_ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->counted) rs_ctx_count(-1);
    delete ctx;
    return CDG_OK;
}

extern "C" const char *cdg_last_error(const cdg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int cdg_ctx_mem_used(const cdg_ctx *ctx, int64_t *bytes) {
    if (!ctx || !bytes) return CDG_ERR_ARG;
    *bytes = ctx->mem_used;
    return CDG_OK;
}

extern "C" int cdg_ctx_stream(const cdg_ctx *ctx, void **stream) {
    if (!ctx || !stream) return CDG_ERR_ARG;
    *stream = static_cast<void *>(ctx->stream);
    return CDG_OK;
}

// Cross-stream ordering without a host wait: an event recorded on one stream,
// waited on by the other.
static int stream_order(cdg_ctx *ctx, hipStream_t before, hipStream_t after) {
    CDG_HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->order_ev) CDG_HIP_TRY(hipEventCreateWithFlags(&ctx->order_ev, hipEventDisableTiming));
    CDG_HIP_TRY(hipEventRecord(ctx->order_ev, before));
    CDG_HIP_TRY(hipStreamWaitEvent(after, ctx->order_ev, 0));
    return CDG_OK;
}

extern "C" int cdg_ctx_wait_stream(cdg_ctx *ctx, void *stream) {
    if (!ctx) return CDG_ERR_ARG;
    return stream_order(ctx, static_cast<hipStream_t>(stream), ctx->stream);
}

extern "C" int cdg_ctx_signal_stream(cdg_ctx *ctx, void *stream) {
    if (!ctx) return CDG_ERR_ARG;
    return stream_order(ctx, ctx->stream, static_cast<hipStream_t>(stream));
}

// ---------------------------------------------------------------------------
// matrix
// ---------------------------------------------------------------------------
extern "C" int cdg_mat_create(cdg_ctx *ctx, int32_t nrows, int32_t block_n, const int32_t *row_ptr,
                              const int32_t *col, const int32_t *kind, cdg_mat **out) {
    if (!ctx || !out) return CDG_ERR_ARG;
    *out = nullptr;
    if (nrows < 0 || block_n <= 0 || block_n > 1024 || !row_ptr || (nrows > 0 && (!col || !kind)))
        return cdg_fail(ctx, CDG_ERR_ARG, "cdg_mat_create: bad sizes or null arrays");
    if (row_ptr[0] != 0) return cdg_fail(ctx, CDG_ERR_ARG, "row_ptr[0] must be 0");
    for (int i = 0; i < nrows; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) return cdg_fail(ctx, CDG_ERR_ARG, "row_ptr not monotone at %d", i);
    const int nnzb = row_ptr[nrows];
    for (int b = 0; b < nnzb; ++b) {
        if (col[b] < 0 || col[b] >= nrows) return cdg_fail(ctx, CDG_ERR_ARG, "col[%d]=%d out of range", b, col[b]);
        if (kind[b] != CDG_KIND_LOCAL && kind[b] != CDG_KIND_COUPLING)
            return cdg_fail(ctx, CDG_ERR_ARG, "kind[%d]=%d invalid", b, kind[b]);
    }
    CDG_HIP_TRY(hipSetDevice(ctx->device));
    cdg_mat *A = new cdg_mat();
    A->nrows = nrows;
    A->n = block_n;
    A->nnzb = nnzb;
    A->ld = slot_elems(block_n);
    A->row_ptr.assign(row_ptr, row_ptr + nrows + 1);
    A->col.assign(col, col + nnzb);
    A->kind.assign(kind, kind + nnzb);
    analyze(A);
    {   // matvec launch shape: tuned for this block size if cdg_mat_tune has run,
        // else ~540 threads with 8 loads in flight each, the measured best at
        // n = 135 (0.611 ms = 6.85 TB/s, against 0.652 ms with ~1024 threads and
        // 4); CDG_MV_THREADS / CDG_MV_UNROLL override either
        MvShape ms = mv_default();
        (void)mv_tuned(block_n, &ms);
        if (const char *t = std::getenv("CDG_MV_THREADS")) ms.target = std::max(64, std::atoi(t));
        if (const char *u = std::getenv("CDG_MV_UNROLL")) ms.unroll = std::atoi(u);
        A->shape64 = stream_shape(block_n, 2, ms.target);
        A->mv_unroll = ms.unroll == 16 ? 16 : ms.unroll == 8 ? 8 : 4;
    }
    int e = dev_alloc(ctx, (void **)&A->blocks, sizeof(double) * (size_t)std::max(1, nnzb) * A->ld);
    if (!e) e = upload_array(ctx, &A->d_row_ptr, A->row_ptr, nullptr);
    if (!e) e = upload_array(ctx, &A->d_col, A->col, nullptr);
    if (e) { cdg_mat_destroy(ctx, A); return e; }
    hipError_t he = hipMemsetAsync(A->blocks, 0, sizeof(double) * (size_t)std::max(1, nnzb) * A->ld, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) { cdg_mat_destroy(ctx, A); return cdg_fail(ctx, CDG_ERR_HIP, "%s", hipGetErrorString(he)); }
    *out = A;
    return CDG_OK;
}

extern "C" int cdg_mat_destroy(cdg_ctx *ctx, cdg_mat *A) {
    if (!A) return CDG_OK;
    if (ctx) {
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(ctx, A->blocks, sizeof(double) * (size_t)std::max(1, A->nnzb) * A->ld);
        dev_free(ctx, A->d_row_ptr, sizeof(int32_t) * (A->row_ptr.size()));
        dev_free(ctx, A->d_col, sizeof(int32_t) * std::max<size_t>(1, A->col.size()));
    }
    delete A;
    return CDG_OK;
}

extern "C" int cdg_mat_upload(cdg_ctx *ctx, cdg_mat *A, const double *blocks) {
    if (!ctx || !A || (!blocks && A->nnzb)) return CDG_ERR_ARG;
    if (A->nnzb == 0) { A->stamp = ++g_stamp; return CDG_OK; }
    const size_t nn = (size_t)A->n * A->n;
    const hipMemcpyKind kind = is_device_ptr(blocks) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    CDG_HIP_TRY(hipMemcpy2DAsync(A->blocks, sizeof(double) * A->ld, blocks, sizeof(double) * nn, sizeof(double) * nn,
                                 A->nnzb, kind, ctx->stream));
    CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    A->stamp = ++g_stamp;
    return CDG_OK;
}

extern "C" int cdg_mat_download(cdg_ctx *ctx, const cdg_mat *A, double *blocks) {
    if (!ctx || !A || (!blocks && A->nnzb)) return CDG_ERR_ARG;
    if (A->nnzb == 0) return CDG_OK;
    const size_t nn = (size_t)A->n * A->n;
    const hipMemcpyKind kind = is_device_ptr(blocks) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    CDG_HIP_TRY(hipMemcpy2DAsync(blocks, sizeof(double) * nn, A->blocks, sizeof(double) * A->ld, sizeof(double) * nn,
                                 A->nnzb, kind, ctx->stream));
    CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return CDG_OK;
}

extern "C" int cdg_mat_fill_synthetic(cdg_ctx *ctx, cdg_mat *A, uint64_t seed, double s_diag, double s_off) {
    if (!ctx || !A) return CDG_ERR_ARG;
    int e = launch_fill_synthetic(ctx, A, seed, s_diag, s_off);
    if (e) return e;
    A->stamp = ++g_stamp;
    return CDG_OK;
}

extern "C" int cdg_vec_fill_synthetic(cdg_ctx *ctx, const cdg_mat *A, uint64_t seed, double scale, double *v) {
    if (!ctx || !A || !v) return CDG_ERR_ARG;
    VecArg o;
    const int64_t len = (int64_t)A->nrows * A->n;
    int e = stage_in(ctx, 2, v, len, o);
    if (!e) e = launch_vec_synthetic(ctx, len, seed, scale, o.dev);
    if (!e) e = finish_out(ctx, o);
    return e;
}

extern "C" int cdg_mat_set_partition(cdg_ctx *ctx, cdg_mat *A, const int32_t *part) {
    if (!ctx || !A) return CDG_ERR_ARG;
    if (part) {
        for (int i = 0; i < A->nrows; ++i)
            if (part[i] < 0) return cdg_fail(ctx, CDG_ERR_ARG, "part[%d] = %d is negative", i, part[i]);
        A->part.assign(part, part + A->nrows);
    } else {
        A->part.clear();
    }
    analyze(A);
    return CDG_OK;
}

extern "C" int cdg_mat_stamp(const cdg_mat *A, uint64_t *stamp) {
    if (!A || !stamp) return CDG_ERR_ARG;
    *stamp = A->stamp;
    return CDG_OK;
}

extern "C" int cdg_mat_info(const cdg_mat *A, int32_t *nrows, int32_t *block_n, int32_t *nnzb) {
    if (!A) return CDG_ERR_ARG;
    if (nrows) *nrows = A->nrows;
    if (block_n) *block_n = A->n;
    if (nnzb) *nnzb = A->nnzb;
    return CDG_OK;
}

extern "C" int cdg_mv(cdg_ctx *ctx, const cdg_mat *A, const double *x, double *y) {
    if (!ctx || !A || !x || !y) return CDG_ERR_ARG;
    const int64_t len = (int64_t)A->nrows * A->n;
    VecArg xi, yo;
    int e = stage_in(ctx, 0, x, len, xi);
    if (!e) e = stage_in(ctx, 1, y, len, yo);
    if (!e) e = upload_in(ctx, xi);
    if (!e) e = launch_mv(ctx, A, xi.dev, yo.dev);
    if (!e) e = finish_out(ctx, yo);
    return e;
}

extern "C" int cdg_mv_async(cdg_ctx *ctx, const cdg_mat *A, const double *x, double *y) {
    if (!ctx || !A || !x || !y) return CDG_ERR_ARG;
    if (!is_device_ptr(x) || !is_device_ptr(y)) return cdg_fail(ctx, CDG_ERR_ARG, "cdg_mv_async needs device pointers");
    return launch_mv(ctx, A, x, y);
}

// ---------------------------------------------------------------------------
// preconditioner
// ---------------------------------------------------------------------------
// Task list, counters and partial storage of the dataflow sweep (once per
// preconditioner).  Tickets: forward levels ascending; per backward level its
// BA tasks then its BB tasks, so every input of a task has a smaller ticket.
static int df_setup(cdg_ctx *ctx, cdg_prec *M) {
    if (M->d_tasks) return CDG_OK;
    int e = df_configure(ctx, M);
    if (e) return e;
    const int S = M->df_S;
    std::vector<int32_t> tasks;
    tasks.reserve((size_t)3 * M->nrows * S);
    for (int l = 0; l < M->nlev; ++l)
        for (int k = M->lev_row_ptr[l]; k < M->lev_row_ptr[l + 1]; ++k)
            for (int sl = 0; sl < S; ++sl) tasks.push_back(((M->h_lev_rows[k] * S + sl) << 2) | 0);
    for (int l = 0; l < M->nblev; ++l)
        for (int stg = 1; stg <= 2; ++stg)
            for (int k = M->blev_row_ptr[l]; k < M->blev_row_ptr[l + 1]; ++k)
                for (int sl = 0; sl < S; ++sl) tasks.push_back(((M->h_blev_rows[k] * S + sl) << 2) | stg);
    M->df_ntasks = (int32_t)tasks.size();
    const int64_t need = 3 * (int64_t)M->nrows * S * M->n;
    if (need > 3 * M->part_len) {      // partial storage shared with the level schedule
        double *np = nullptr;
        e = dev_alloc(ctx, (void **)&np, sizeof(double) * (size_t)(need + 2));   // +2 as at init
        if (e) return e;
        for (DevMem &m : M->owned)
            if (m.p == M->d_part) { dev_free(ctx, m.p, m.bytes); m.p = np; m.bytes = sizeof(double) * (size_t)(need + 2); }
        M->d_part = np;
        M->part_len = need / 3;
        M->part_cap = need + 2;
    }
    e = upload_array(ctx, &M->d_tasks, tasks, &M->owned);
    if (!e) e = upload_array(ctx, &M->d_cnt, std::vector<int32_t>((size_t)3 * M->nrows, 0), &M->owned);
    if (!e) e = upload_array(ctx, &M->d_dferr, std::vector<int32_t>(1, 0), &M->owned);
    if (!e) {
        e = dev_alloc(ctx, (void **)&M->d_ticket, sizeof(unsigned long long));
        if (!e) M->owned.push_back({M->d_ticket, sizeof(unsigned long long)});
    }
    if (!e) e = df_reset(ctx, M);
    return e;
}

extern "C" int cdg_prec_destroy(cdg_ctx *ctx, cdg_prec *M) {
    if (!M) return CDG_OK;
    if (ctx) {
        (void)hipStreamSynchronize(ctx->stream);
        for (DevMem &m : M->owned) dev_free(ctx, m.p, m.bytes);
        if (M->ym) dev_free(ctx, M->ym, M->ym_bytes);
    }
    delete M;
    return CDG_OK;
}

extern "C" int cdg_prec_create(cdg_ctx *ctx, const cdg_mat *A, int32_t type, int32_t precision, cdg_prec **out) {
    if (!ctx || !A || !out) return CDG_ERR_ARG;
    *out = nullptr;
    if (type != CDG_PREC_IDENTITY && type != CDG_PREC_ILU0 && type != CDG_PREC_JACOBI)
        return cdg_fail(ctx, CDG_ERR_ARG, "unknown preconditioner %d", type);
    if (precision < CDG_FP64 || precision > CDG_BF16) return cdg_fail(ctx, CDG_ERR_ARG, "unknown precision %d", precision);
    cdg_prec *M = new cdg_prec();
    M->type = type;
    M->precision = precision;
    M->nrows = A->nrows;
    M->n = A->n;
    if (type == CDG_PREC_IDENTITY) { *out = M; return CDG_OK; }
    // The kind-0 structure the factor follows: A's, or for block Jacobi
    // (precon_jacobi.f90:95-215) the diagonal alone -- ILU(0) of a matrix
    // whose every row is its own rank: L = 0, U = 0, Dinv_i = inv(A_ii), and
    // the apply is z_i = Dinv_i v_i.
    cdg_mat jview;
    const cdg_mat *S = A;
    if (type == CDG_PREC_JACOBI) {
        jview.nrows = A->nrows;
        jview.n = A->n;
        jview.nnzb = A->nnzb;
        jview.row_ptr = A->row_ptr;
        jview.col = A->col;
        jview.kind = A->kind;
        jview.part.resize(A->nrows);
        for (int i = 0; i < A->nrows; ++i) jview.part[i] = i;
        analyze(&jview);
        S = &jview;
        M->type = CDG_PREC_ILU0;   // the ILU(0) machinery with no off-diagonal blocks
        M->jacobi = 1;
    }
    if (S->struct_err != CDG_OK) {
        delete M;
        return cdg_fail(ctx, S->struct_err, S->struct_err == CDG_ERR_NO_TRANSPOSE
                                                ? "lblks entry without a transposed partner"
                                                : "block row without a diagonal block");
    }
    M->nlo = (int32_t)S->lo_blk.size();
    M->nup = (int32_t)S->up_blk.size();
    M->ld = slot_elems(A->n);
    const size_t es = precision == CDG_FP64 ? 8 : (precision == CDG_FP32 ? 4 : 2);
    auto alloc = [&](void **p, size_t bytes) -> int {
        bytes = std::max<size_t>(bytes, 16);
        int e = dev_alloc(ctx, p, bytes);
        if (!e) M->owned.push_back({*p, bytes});
        return e;
    };
    int e = alloc(&M->L, es * (size_t)M->nlo * M->ld);
    if (!e) e = alloc(&M->Dinv, es * (size_t)M->nrows * M->ld);
    if (!e && precision != CDG_FP64) e = alloc(&M->U, es * (size_t)M->nup * M->ld);
    if (!e && precision == CDG_BF16) {
        e = alloc((void **)&M->L32, 4 * (size_t)M->nlo * M->ld);
        if (!e) e = alloc((void **)&M->D32, 4 * (size_t)M->nrows * M->ld);
    }
    // structure arrays
    std::vector<int32_t> up_idx(M->nup);
    for (int u = 0; u < M->nup; ++u) up_idx[u] = precision == CDG_FP64 ? S->up_blk[u] : u;
    if (!e) e = upload_array(ctx, &M->d_lo_ptr, S->lo_ptr, &M->owned);
    if (!e) e = upload_array(ctx, &M->d_lo_blk, S->lo_blk, &M->owned);
    if (!e) e = upload_array(ctx, &M->d_lo_col, S->lo_col, &M->owned);
    if (!e) e = upload_array(ctx, &M->d_lo_tr, S->lo_tr, &M->owned);
    if (!e && precision != CDG_FP64) {
        // A_ji (j < i) is an upper block of row j, so every lo_tr[q] has a U slot
        std::vector<int32_t> slot_of(std::max(A->nnzb, 1), -1), lo_tru(S->lo_tr.size());
        for (int u = 0; u < M->nup; ++u) slot_of[S->up_blk[u]] = u;
        for (size_t q = 0; q < lo_tru.size(); ++q) lo_tru[q] = slot_of[S->lo_tr[q]];
        e = upload_array(ctx, &M->d_lo_tru, lo_tru, &M->owned);
    }
    if (!e) e = upload_array(ctx, &M->d_up_ptr, S->up_ptr, &M->owned);
    if (!e) e = upload_array(ctx, &M->d_up_blk, S->up_blk, &M->owned);
    if (!e) e = upload_array(ctx, &M->d_up_idx, up_idx, &M->owned);
    if (!e) e = upload_array(ctx, &M->d_up_col, S->up_col, &M->owned);
    if (!e) e = upload_array(ctx, &M->d_diag, S->diag, &M->owned);
    // forward levels (factor schedule and forward sweep), backward levels
    std::vector<int32_t> fptr, bptr;
    std::vector<int32_t> forder = order_by_level(S->flevel, S->nflev, &fptr);
    std::vector<int32_t> border = order_by_level(S->blevel, S->nblev, &bptr);
    M->nlev = S->nflev;
    M->lev_row_ptr = fptr;
    M->nblev = S->nblev;
    M->blev_row_ptr = bptr;
    std::vector<int32_t> lev_lo;
    M->lev_lo_ptr.assign(M->nlev + 1, 0);
    for (int l = 0; l < M->nlev; ++l) {
        for (int k = fptr[l]; k < fptr[l + 1]; ++k) {
            const int i = forder[k];
            for (int q = S->lo_ptr[i]; q < S->lo_ptr[i + 1]; ++q) lev_lo.push_back(q);
        }
        M->lev_lo_ptr[l + 1] = (int32_t)lev_lo.size();
    }
    if (!e) e = upload_array(ctx, &M->d_lev_lo, lev_lo, &M->owned);
    {   // the W jobs: every upper slot
        std::vector<int32_t> all(std::max(M->nup, 1));
        for (int u = 0; u < M->nup; ++u) all[u] = u;
        if (!e) e = upload_array(ctx, &M->d_up_all, all, &M->owned);
    }
    if (!e) e = upload_array(ctx, &M->d_lev_rows, forder, &M->owned);
    if (!e) e = upload_array(ctx, &M->d_blev_rows, border, &M->owned);
    if (!e) e = upload_array(ctx, &M->d_err, std::vector<int32_t>(1, 0), &M->owned);
    if (!e) {
        hipError_t he = hipStreamSynchronize(ctx->stream);
        if (he != hipSuccess) e = cdg_fail(ctx, CDG_ERR_HIP, "%s", hipGetErrorString(he));
    }
    if (e) { cdg_prec_destroy(ctx, M); return e; }
    M->maxdep = 1;
    for (int i = 0; i < A->nrows; ++i) {
        M->maxdep = std::max(M->maxdep, S->lo_ptr[i + 1] - S->lo_ptr[i]);
        M->maxdep = std::max(M->maxdep, S->up_ptr[i + 1] - S->up_ptr[i]);
    }
    // Apply launch shape: the tuned one for this block size and precision if
    // cdg_prec_tune has run, else the default; CDG_APPLY_SLICE_W /
    // CDG_APPLY_THREADS / CDG_APPLY_UNROLL override either.
    ApplyShape ash = apply_default(A->n, precision);
    (void)apply_tuned(A->n, precision, &ash);
    if (const char *ev = std::getenv("CDG_APPLY_SLICE_W")) ash.sw = std::max(8, std::atoi(ev));
    if (const char *ev = std::getenv("CDG_APPLY_THREADS")) ash.target = std::max(64, std::atoi(ev));
    if (const char *ev = std::getenv("CDG_APPLY_UNROLL")) ash.unroll = std::atoi(ev) == 8 ? 8 : 4;
    if (e || !apply_shape_set(M, precision, ash)) {
        cdg_prec_destroy(ctx, M);
        return e ? e : cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "ILU apply launch shape does not fit (block size %d)", A->n);
    }
    // apply schedule (CDG_APPLY_MODE=levels|dataflow|folded|auto overrides the default)
    M->apply_mode = CDG_APPLY_AUTO;
    if (const char *m = std::getenv("CDG_APPLY_MODE")) {
        if (!std::strcmp(m, "dataflow")) M->apply_mode = CDG_APPLY_DATAFLOW;
        else if (!std::strcmp(m, "auto")) M->apply_mode = CDG_APPLY_AUTO;
        else if (!std::strcmp(m, "levels")) M->apply_mode = CDG_APPLY_LEVELS;
        else if (!std::strcmp(m, "folded") && precision != CDG_BF16) M->apply_mode = CDG_APPLY_FOLDED;
        else if (!std::strcmp(m, "stream")) M->apply_mode = CDG_APPLY_STREAM;
    }
    M->h_lev_rows = forder;
    M->h_lo_ptr = S->lo_ptr;
    M->h_lo_col = S->lo_col;
    M->h_up_ptr = S->up_ptr;
    M->h_up_col = S->up_col;
    M->h_blev_rows = border;
    // (+2: a sweep launch whose rows have no entries issues its 16-byte
    // block prefetch from the start of a partial array, unused)
    const int64_t part_doubles = 3 * M->part_len + 2;
    if (!e) e = alloc((void **)&M->d_part, sizeof(double) * (size_t)part_doubles);
    M->part_cap = part_doubles;
    if (!e && M->apply_mode == CDG_APPLY_DATAFLOW) e = df_setup(ctx, M);
    if (!e && M->apply_mode == CDG_APPLY_FOLDED) e = alloc(&M->W, es * (size_t)M->nup * M->ld);
    if (!e && M->apply_mode == CDG_APPLY_STREAM) e = rs_setup(ctx, M);
    if (const char *f = std::getenv("CDG_FOLD_AFTER")) M->fold_after = std::max(0, std::atoi(f));
    if (!e) {
        hipError_t he = hipStreamSynchronize(ctx->stream);
        if (he != hipSuccess) e = cdg_fail(ctx, CDG_ERR_HIP, "%s", hipGetErrorString(he));
    }
    if (e) { cdg_prec_destroy(ctx, M); return e; }
    *out = M;
    return CDG_OK;
}

extern "C" int cdg_prec_set_schedule(cdg_ctx *ctx, cdg_prec *M, int32_t schedule) {
    if (!ctx || !M) return CDG_ERR_ARG;
    if (schedule != CDG_SCHEDULE_LEVELS && schedule != CDG_SCHEDULE_DATAFLOW && schedule != CDG_SCHEDULE_FOLDED &&
        schedule != CDG_SCHEDULE_AUTO && schedule != CDG_SCHEDULE_STREAM && schedule != CDG_SCHEDULE_ROWSLICE)
        return cdg_fail(ctx, CDG_ERR_ARG, "unknown apply schedule %d", schedule);
    if (M->type != CDG_PREC_ILU0) return CDG_OK;
    if (schedule == CDG_SCHEDULE_DATAFLOW) {
        int e = df_setup(ctx, M);
        if (e) return e;
        CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    if (schedule == CDG_SCHEDULE_STREAM || schedule == CDG_SCHEDULE_ROWSLICE) {   // storage now; W and the copies at the next apply
        int e = rs_setup(ctx, M);
        if (e) return e;
        CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    if (schedule == CDG_SCHEDULE_FOLDED) {
        if (M->precision == CDG_BF16)
            return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "folded schedule: FP64 and FP32 factors only");
        if (!M->W) {
            const size_t bytes = std::max<size_t>(16, (M->precision == CDG_FP64 ? 8 : 4) * (size_t)M->nup * M->ld);
            int e = dev_alloc(ctx, &M->W, bytes);
            if (e) return e;
            M->owned.push_back({M->W, bytes});
        }
    }
    M->apply_mode = schedule == CDG_SCHEDULE_DATAFLOW ? CDG_APPLY_DATAFLOW
                    : schedule == CDG_SCHEDULE_FOLDED ? CDG_APPLY_FOLDED
                    : schedule == CDG_SCHEDULE_STREAM ? CDG_APPLY_STREAM
                    : schedule == CDG_SCHEDULE_ROWSLICE ? CDG_APPLY_ROWSLICE
                    : schedule == CDG_SCHEDULE_AUTO ? CDG_APPLY_AUTO : CDG_APPLY_LEVELS;
    return CDG_OK;
}

// ---- launch-shape autotuner (no reference counterpart) ---------------------------
namespace {
// Times fn() `reps` times on the context's stream (one untimed call first).
template <typename F>
int time_on_stream(cdg_ctx *ctx, hipEvent_t e0, hipEvent_t e1, int reps, F fn, double *ms_out) {
    int e = fn();
    if (e) return e;
    CDG_HIP_TRY(hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < reps && !e; ++r) e = fn();
    if (e) return e;
    CDG_HIP_TRY(hipEventRecord(e1, ctx->stream));
    CDG_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    CDG_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    *ms_out = (double)ms / reps;
    return CDG_OK;
}

// Grows the apply's partial-sum storage to the current part_len.
int prec_part_fit(cdg_ctx *ctx, cdg_prec *M) {
    const int64_t need = 3 * M->part_len + 2;
    if (need <= M->part_cap) return CDG_OK;
    double *np = nullptr;
    int e = dev_alloc(ctx, (void **)&np, sizeof(double) * (size_t)need);
    if (e) return e;
    for (DevMem &m : M->owned)
        if (m.p == M->d_part) { dev_free(ctx, m.p, m.bytes); m.p = np; m.bytes = sizeof(double) * (size_t)need; }
    M->d_part = np;
    M->part_cap = need;
    return CDG_OK;
}

struct Scratch {   // tuning vectors and events, released on every path out
    cdg_ctx *ctx;
    double *v = nullptr, *z = nullptr;
    size_t bytes = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    explicit Scratch(cdg_ctx *c) : ctx(c) {}
    int init(int64_t len) {
        bytes = sizeof(double) * (size_t)std::max<int64_t>(len, 1);
        int e = dev_alloc(ctx, (void **)&v, bytes);
        if (!e) e = dev_alloc(ctx, (void **)&z, bytes);
        if (!e) e = kv_zero(ctx, std::max<int64_t>(len, 1), v);
        if (!e && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
            e = cdg_fail(ctx, CDG_ERR_HIP, "hipEventCreate failed");
        return e;
    }
    ~Scratch() {
        (void)hipStreamSynchronize(ctx->stream);
        if (v) dev_free(ctx, v, bytes);
        if (z) dev_free(ctx, z, bytes);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};
}  // namespace

extern "C" int cdg_prec_tune(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, int32_t reps, double *best_ms) {
    if (!ctx || !M || !A) return CDG_ERR_ARG;
    if (best_ms) *best_ms = 0.0;
    if (M->type != CDG_PREC_ILU0 || M->nrows == 0) return CDG_OK;
    if (A->nrows != M->nrows || A->n != M->n) return cdg_fail(ctx, CDG_ERR_ARG, "matrix does not match preconditioner");
    if (!M->factored) return cdg_fail(ctx, CDG_ERR_STATE, "cdg_prec_tune before cdg_prec_update");
    if (M->apply_mode == CDG_APPLY_DATAFLOW)
        return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "cdg_prec_tune: the level schedules only");
    reps = std::max(1, std::min<int32_t>(reps, 1000));
    const int n = M->n, prec = M->precision;
    Scratch sc(ctx);
    int e = sc.init((int64_t)M->nrows * n);
    if (e) return e;
    // the default first; another shape must beat the best so far by 2 % (the
    // run-to-run spread of one candidate's timing) to replace it
    const ApplyShape def = apply_default(n, prec);
    std::vector<ApplyShape> cands{def};
    for (int sw : {16, 24, 32, 40, 48, 64, 96, 128, n})
        for (int tg : {128, 256, 512})
            for (int un : {4, 8}) cands.push_back({sw, tg, un});
    std::vector<std::array<int, 4>> seen;
    double best = std::numeric_limits<double>::max();
    ApplyShape bestc = def;
    for (const ApplyShape &c : cands) {
        if (!apply_shape_set(M, prec, c)) continue;
        const std::array<int, 4> key{M->slices, M->slice_w, M->shape[prec].T, M->unroll};
        if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
        seen.push_back(key);
        e = prec_part_fit(ctx, M);
        double ms = 0.0;
        if (!e) e = time_on_stream(ctx, sc.e0, sc.e1, reps,
                                   [&] { return launch_ilu_apply(ctx, M, A, sc.v, sc.z); }, &ms);
        if (e) {
            (void)apply_shape_set(M, prec, bestc);
            return e;
        }
        if (ms < 0.98 * best || best == std::numeric_limits<double>::max()) { best = ms; bestc = c; }
    }
    if (!apply_shape_set(M, prec, bestc)) return cdg_fail(ctx, CDG_ERR_STATE, "tuned shape no longer fits");
    e = prec_part_fit(ctx, M);
    if (e) return e;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_apply_tuned[{n, prec}] = bestc;
    }
    if (best_ms) *best_ms = best;
    return CDG_OK;
}

extern "C" int cdg_mat_tune(cdg_ctx *ctx, cdg_mat *A, int32_t reps, double *best_ms) {
    if (!ctx || !A) return CDG_ERR_ARG;
    if (best_ms) *best_ms = 0.0;
    if (A->nrows == 0) return CDG_OK;
    reps = std::max(1, std::min<int32_t>(reps, 1000));
    Scratch sc(ctx);
    int e = sc.init((int64_t)A->nrows * A->n);
    if (e) return e;
    // the default first; another shape must beat the best so far by 2 %
    const MvShape def = mv_default();
    std::vector<MvShape> cands{def};
    for (int tg : {128, 256, 384, 540, 768, 1024})
        for (int un : {4, 8, 16}) cands.push_back({tg, un});
    std::vector<std::array<int, 2>> seen;
    double best = std::numeric_limits<double>::max();
    const StreamShape keep = A->shape64;
    const int32_t keep_u = A->mv_unroll;
    MvShape bestc{0, keep_u};
    StreamShape bests = keep;
    for (const MvShape &cand : cands) {
        const int tg = cand.target, un = cand.unroll;
        const StreamShape sh = stream_shape(A->n, 2, tg);
        if (sh.T == 0) continue;
        const std::array<int, 2> key{sh.T, un};
        if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
        seen.push_back(key);
        A->shape64 = sh;
        A->mv_unroll = un;
        double ms = 0.0;
        e = time_on_stream(ctx, sc.e0, sc.e1, reps, [&] { return launch_mv(ctx, A, sc.v, sc.z); }, &ms);
        if (e) {
            A->shape64 = keep;
            A->mv_unroll = keep_u;
            return e;
        }
        if (ms < 0.98 * best || best == std::numeric_limits<double>::max()) { best = ms; bestc = {tg, un}; bests = sh; }
    }
    A->shape64 = bests;
    A->mv_unroll = bestc.unroll;
    if (bestc.target > 0) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_mv_tuned[A->n] = bestc;
    }
    if (best_ms) *best_ms = best;
    return CDG_OK;
}

extern "C" int cdg_prec_launch_shape(const cdg_prec *M, int32_t *out) {
    if (!M || !out) return CDG_ERR_ARG;
    out[0] = M->slices;
    out[1] = M->slice_w;
    out[2] = M->shape[M->precision].T;
    out[3] = M->unroll;
    return CDG_OK;
}

extern "C" int cdg_mat_launch_shape(const cdg_mat *A, int32_t *out) {
    if (!A || !out) return CDG_ERR_ARG;
    out[0] = A->shape64.T;
    out[1] = A->mv_unroll;
    return CDG_OK;
}

// A dataflow sweep that gave up waiting leaves its counters unusable: report
// it and reset them (called from the synchronizing entry points).
static int df_check(cdg_ctx *ctx, cdg_prec *M) {
    if (M->d_rs_ctl) {   // the stream sweep: an abandoned wait left NaN in z
        uint32_t err = 0;
        CDG_HIP_TRY(hipMemcpyAsync(&err, M->d_rs_ctl, sizeof(err), hipMemcpyDeviceToHost, ctx->stream));
        CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (err) {
            CDG_HIP_TRY(hipMemsetAsync(M->d_rs_ctl, 0, sizeof(uint32_t), ctx->stream));
            return cdg_fail(ctx, CDG_ERR_HIP, "ILU(0) stream sweep gave up waiting for a dependency");
        }
    }
    if (M->apply_mode != CDG_APPLY_DATAFLOW || !M->d_dferr) return CDG_OK;
    int32_t err = 0;
    CDG_HIP_TRY(hipMemcpyAsync(&err, M->d_dferr, sizeof(err), hipMemcpyDeviceToHost, ctx->stream));
    CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (!err) return CDG_OK;
    int e = df_reset(ctx, M);
    if (e) return e;
    return cdg_fail(ctx, CDG_ERR_HIP, "ILU(0) dataflow sweep gave up waiting for a dependency; counters reset");
}

extern "C" int cdg_prec_update(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, int32_t *bad_row) {
    if (!ctx || !M || !A) return CDG_ERR_ARG;
    if (bad_row) *bad_row = -1;
    if (A->nrows != M->nrows || A->n != M->n) return cdg_fail(ctx, CDG_ERR_ARG, "matrix does not match preconditioner");
    if (M->type == CDG_PREC_IDENTITY) { M->stamp = A->stamp; M->factored = 1; return CDG_OK; }
    M->factored = 0;
    M->w_ready = 0;
    M->rs_ready = 0;
    M->prev_napply = M->napply;
    M->napply = 0;
    int e = launch_ilu_factor(ctx, M, A, bad_row);
    if (e) return e;
    M->stamp = A->stamp;
    M->factored = 1;
    return CDG_OK;
}

// The stream schedule's storage for the current factor: W = Dinv U and the
// row-slice copies of L and Dinv (once per factor).
int rs_prepare(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A) {
    if (M->rs_ready) return CDG_OK;
    int e = rs_setup(ctx, M);
    if (!e) e = launch_fold_rs(ctx, M, A);
    if (!e) e = rs_relayout(ctx, M);
    if (!e) M->rs_ready = 1;
    return e;
}

static int prec_apply_dev(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    const int64_t len = (int64_t)M->nrows * M->n;
    if (M->type == CDG_PREC_IDENTITY) return kv_copy(ctx, len, v, z);
    if (!M->factored) return cdg_fail(ctx, CDG_ERR_STATE, "ILU0 apply before update");
    if (M->apply_mode == CDG_APPLY_DATAFLOW) return launch_ilu_apply_df(ctx, M, A, v, z);
    ++M->napply;
    const bool auto_long = M->apply_mode == CDG_APPLY_AUTO && M->fold_after > 0 &&
                           (M->napply > M->fold_after || M->prev_napply > M->fold_after) && M->rs_ok;
    if (M->apply_mode == CDG_APPLY_ROWSLICE || M->apply_mode == CDG_APPLY_STREAM || auto_long) {
        int e = rs_prepare(ctx, M, A);
        if (e && M->apply_mode == CDG_APPLY_AUTO && e == CDG_ERR_UNSUPPORTED) {
            M->rs_ok = 0;   // this pattern or block size: the level schedule from now on
        } else {
            if (e) return e;
            // AUTO settles in the row-slice level launches (measured faster
            // than the persistent sweep on the same storage)
            return M->apply_mode == CDG_APPLY_STREAM ? launch_ilu_apply_rs(ctx, M, v, z)
                                                     : launch_ilu_apply_rsl(ctx, M, v, z);
        }
    }
    const bool fold_now = !M->w_ready && M->apply_mode == CDG_APPLY_FOLDED;   // schedule chosen after the factor
    if (fold_now) {
        if (!M->W) {
            const size_t bytes = std::max<size_t>(16, (M->precision == CDG_FP64 ? 8 : 4) * (size_t)M->nup * M->ld);
            int e = dev_alloc(ctx, &M->W, bytes);
            if (e) return e;
            M->owned.push_back({M->W, bytes});
        }
        int e = launch_fold_all(ctx, M, A);
        if (e) return e;
        M->w_ready = 1;
    }
    return launch_ilu_apply(ctx, M, A, v, z);
}

extern "C" int cdg_prec_apply(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    if (!ctx || !M || !A || !v || !z) return CDG_ERR_ARG;
    const int64_t len = (int64_t)M->nrows * M->n;
    VecArg vi, zo;
    int e = stage_in(ctx, 0, v, len, vi);
    if (!e) e = stage_in(ctx, 1, z, len, zo);
    if (!e) e = upload_in(ctx, vi);
    if (!e) e = prec_apply_dev(ctx, M, A, vi.dev, zo.dev);
    if (!e) e = finish_out(ctx, zo);
    if (!e) e = df_check(ctx, M);
    return e;
}

extern "C" int cdg_prec_apply_async(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    if (!ctx || !M || !A || !v || !z) return CDG_ERR_ARG;
    if (!is_device_ptr(v) || !is_device_ptr(z)) return cdg_fail(ctx, CDG_ERR_ARG, "cdg_prec_apply_async needs device pointers");
    return prec_apply_dev(ctx, M, A, v, z);
}

extern "C" int cdg_prec_download(cdg_ctx *ctx, const cdg_prec *M, double *L, double *Dinv) {
    if (!ctx || !M) return CDG_ERR_ARG;
    if (M->type != CDG_PREC_ILU0 || !M->factored) return cdg_fail(ctx, CDG_ERR_STATE, "no ILU0 factor to download");
    const size_t nn = (size_t)M->n * M->n;
    double *tmp = nullptr;
    const int64_t cnt = (int64_t)std::max(M->nlo, M->nrows) * M->ld;
    int e = dev_alloc(ctx, (void **)&tmp, sizeof(double) * std::max<int64_t>(cnt, 1));
    if (e) return e;
    std::vector<double> h((size_t)std::max<int64_t>(cnt, 1));
    const size_t es = M->precision == CDG_FP64 ? 8 : (M->precision == CDG_FP32 ? 4 : 2);
    auto fetch = [&](const void *src, int nslots, auto put) -> int {
        if (nslots == 0) return CDG_OK;
        int ee = launch_widen(ctx, src, M->precision, tmp, (int64_t)nslots * M->ld);
        if (ee) return ee;
        CDG_HIP_TRY(hipMemcpyAsync(h.data(), tmp, sizeof(double) * (size_t)nslots * M->ld, hipMemcpyDeviceToHost, ctx->stream));
        CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (int s = 0; s < nslots; ++s) put(s, h.data() + (size_t)s * M->ld);
        return CDG_OK;
    };
    (void)es;
    // slot q of L belongs to the q-th lower block in pattern order
    std::vector<int32_t> lo_blk_host;
    // The prec does not keep the pattern; recover block indices from the device copy.
    if (M->nlo) {
        lo_blk_host.resize(M->nlo);
        hipError_t he = hipMemcpy(lo_blk_host.data(), M->d_lo_blk, sizeof(int32_t) * M->nlo, hipMemcpyDeviceToHost);
        if (he != hipSuccess) { dev_free(ctx, tmp, sizeof(double) * cnt); return cdg_fail(ctx, CDG_ERR_HIP, "%s", hipGetErrorString(he)); }
    }
    if (L) {
        e = fetch(M->L, M->nlo, [&](int q, const double *src) {
            std::memcpy(L + (size_t)lo_blk_host[q] * nn, src, sizeof(double) * nn);
        });
    }
    if (!e && Dinv) {
        e = fetch(M->Dinv, M->nrows, [&](int i, const double *src) {
            std::memcpy(Dinv + (size_t)i * nn, src, sizeof(double) * nn);
        });
    }
    dev_free(ctx, tmp, sizeof(double) * std::max<int64_t>(cnt, 1));
    return e;
}

// ---------------------------------------------------------------------------
// FGMRES (type_fgmres.f90:81-310)
// ---------------------------------------------------------------------------
extern "C" void cdg_fgmres_default_params(cdg_fgmres_params *p) {
    if (!p) return;
    p->nkrylov = 2000;
    p->tol = 1.e-8;
    p->rtol = 1.e-10;
    p->nmax = -1;
    p->orthogonalization = CDG_CGS;
    p->inner_fgmres = 1;
    p->inner_nkrylov = 100;
    p->inner_tol = 1.e-1;
    p->inner_rtol = 1.e-1;
    p->inner_nmax = 100;
    p->inner_orthogonalization = CDG_CGS;
    p->update_preconditioner = 1;
}

namespace {

struct Workspace {
    int m = 0;
    int64_t len = 0;
    double *V = nullptr, *Z = nullptr, *w = nullptr, *r0 = nullptr, *x0 = nullptr, *zr = nullptr, *dz = nullptr;
    size_t bytes = 0;
    void *base = nullptr;
};

int ws_alloc(cdg_ctx *ctx, Workspace &ws, int m, int64_t len, bool inner_bufs) {
    ws.m = m;
    ws.len = len;
    const size_t lv = (size_t)std::max<int64_t>(len, 1);
    const size_t nvec = 2 * (size_t)(m + 1) + 3 + (inner_bufs ? 2 : 0);
    ws.bytes = nvec * lv * sizeof(double);
    int e = dev_alloc(ctx, &ws.base, ws.bytes);
    if (e) return e;
    double *p = static_cast<double *>(ws.base);
    ws.V = p; p += (size_t)(m + 1) * lv;
    ws.Z = p; p += (size_t)(m + 1) * lv;
    ws.w = p; p += lv;
    ws.r0 = p; p += lv;
    ws.x0 = p; p += lv;
    if (inner_bufs) { ws.zr = p; p += lv; ws.dz = p; p += lv; }
    return CDG_OK;
}

void ws_free(cdg_ctx *ctx, Workspace &ws) {
    dev_free(ctx, ws.base, ws.bytes);
    ws = Workspace();
}

struct Timers {
    double precon = 0, mv = 0, dot = 0, norm = 0;
};

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Solver {
    cdg_ctx *ctx;
    const cdg_mat *A;
    cdg_prec *M;
    int64_t len;
    Timers t;
    int nreorth = 0, nrestart = 0;
    int status = CDG_OK;
    // per nesting level: events before the preconditioner, after it, after
    // the matvec (timing without synchronizing the stream)
    hipEvent_t ev[2][3] = {};
};

// fetch n device scalars from ctx->d_small into ctx->h_small
int fetch_small(cdg_ctx *ctx, size_t n) {
    CDG_HIP_TRY(hipMemcpyAsync(ctx->h_small, ctx->d_small, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    return ctx_wait(ctx);
}

// Orthogonalization of w = V_jj (the slot of the next basis vector) against
// V_0..V_{jj-1}: host hcol += the products, *nrm2 = ||w||^2 after, and with
// normb2 also ||w||^2 before (CGS: w . w is one more product of the same
// launch).  With `reorth` (CGS, first pass) the selective re-orthogonalization
// test runs on the device and its second pass follows at once, gated by the
// test's flag, so both passes cost one device-to-host fetch; *reorth tells
// whether the second pass ran.  Otherwise one pass (*reorth = -1 when given:
// the caller applies the test).
// d_small layout from d_h: [0, jj) products, [jj] ||w||^2 before, [jj+1]
// after; with the gated pass also [jj+2] flag, [jj+3, 2jj+3) second
// products, [2jj+3] ||w||^2 after the second pass.
int orthogonalize(Solver &S, Workspace &ws, int orth, int jj, double *w, double *hcol, double *nrm2,
                  double *normb2, int *reorth = nullptr) {
    cdg_ctx *ctx = S.ctx;
    const int64_t len = S.len;
    if (w != ws.V + (size_t)jj * len) return cdg_fail(ctx, CDG_ERR_STATE, "orthogonalize: w is not V_%d", jj);
    int e = ctx_ensure_small(ctx, 2 * (size_t)jj + 4 + 8 + ws.m + 8);
    if (e) return e;
    double *d_h = ctx->d_small + 8 + ws.m + 8;   // beyond the combine scratch
    const bool gated = reorth && normb2 && orth != CDG_MGS;
    if (orth == CDG_MGS) {
        if (normb2) e = kv_sumsq(ctx, len, w, d_h + jj);
        for (int i = 0; i < jj && !e; ++i) {
            e = kv_multidot(ctx, len, w, ws.V + (size_t)i * len, len, 1, d_h + i);
            if (!e) e = kv_axpy_dev(ctx, len, w, ws.V + (size_t)i * len, d_h + i);
        }
        if (!e) e = kv_sumsq(ctx, len, w, d_h + jj + 1);
    } else {
        // the gated pass's flag comes with the first update's norm (the same launches)
        double *flag = gated ? d_h + jj + 2 : nullptr, *h2 = d_h + jj + 3;
        e = kv_multidot(ctx, len, w, ws.V, len, normb2 ? jj + 1 : jj, d_h);
        if (!e) e = kv_cgs_update(ctx, len, w, ws.V, len, jj, d_h, d_h + jj + 1, nullptr, flag);
        if (gated) {
            if (!e) e = kv_multidot(ctx, len, w, ws.V, len, jj, h2, flag);
            if (!e) e = kv_cgs_update(ctx, len, w, ws.V, len, jj, h2, h2 + jj, flag);
        }
    }
    if (e) return e;
    const size_t nfetch = gated ? 2 * (size_t)jj + 4 : (size_t)jj + 2;
    CDG_HIP_TRY(hipMemcpyAsync(ctx->h_small, d_h, sizeof(double) * nfetch, hipMemcpyDeviceToHost, ctx->stream));
    e = ctx_wait(ctx);
    if (e) return e;
    const double *hs = ctx->h_small;
    for (int i = 0; i < jj; ++i) hcol[i] += hs[i];
    if (normb2) *normb2 = hs[jj];
    *nrm2 = hs[jj + 1];
    if (reorth) *reorth = -1;
    if (gated) {
        *reorth = hs[jj + 2] != 0.0;
        if (*reorth) {
            for (int i = 0; i < jj; ++i) hcol[i] += hs[jj + 3 + i];
            *nrm2 = hs[2 * jj + 3];
        }
    }
    return CDG_OK;
}

int fgmres_core(Solver &S, std::vector<Workspace> &wss, int depth, const double *b, double *x, int m, double tol,
                double rtol, int nmax, int orth, const cdg_fgmres_params *inner, int *niter_out) {
    cdg_ctx *ctx = S.ctx;
    const int64_t len = S.len;
    if (m < 1) return cdg_fail(ctx, CDG_ERR_ARG, "nkrylov must be >= 1");
    if (depth >= (int)wss.size()) return cdg_fail(ctx, CDG_ERR_ARG, "FGMRES nesting deeper than one inner level");
    Workspace &ws = wss[depth];   // wss is never resized during a solve
    const bool need_inner = inner != nullptr;
    if (ws.m != m || ws.len != len || (need_inner && !ws.zr)) {
        ws_free(ctx, ws);
        int e = ws_alloc(ctx, ws, m, len, need_inner);
        if (e) return e;
    }
    std::vector<double> h((size_t)(m + 1) * m), c(m + 1), s(m + 1), p(m + 1), y(m + 1);
    auto H = [&](int i, int j) -> double & { return h[(size_t)j * (m + 1) + i]; };
    int e = kv_zero(ctx, len, ws.x0);
    if (!e) e = kv_zero(ctx, len, x);
    if (e) return e;
    int niter = 0;
    double res = std::numeric_limits<double>::max(), r0norm_initial = 0.0;
    bool conv = false;
    while (res > tol) {
        std::fill(h.begin(), h.end(), 0.0);
        std::fill(p.begin(), p.end(), 0.0);
        std::fill(c.begin(), c.end(), 0.0);
        std::fill(s.begin(), s.end(), 0.0);
        // r0 = b - A x0 ; v1 = r0/||r0|| ; p1 = ||r0||   (:160-163)
        auto t0 = clk::now();
        e = launch_mv(ctx, S.A, ws.x0, ws.r0);
        if (!e) e = kv_sub(ctx, len, b, ws.r0, ws.r0);
        if (!e) e = kv_sumsq(ctx, len, ws.r0, ctx->d_small);
        if (!e) e = fetch_small(ctx, 1);
        if (e) return e;
        const double r0norm = std::sqrt(ctx->h_small[0]);
        S.t.norm += secs(t0, clk::now());
        if (!std::isfinite(r0norm)) return cdg_fail(ctx, CDG_ERR_NAN, "non-finite residual norm");
        if (niter == 0) r0norm_initial = r0norm;
        if (r0norm == 0.0) {
            // exact x0: the reference divides by zero here (type_fgmres.f90:162); return x = x0
            e = kv_copy(ctx, len, ws.x0, x);
            if (e) return e;
            res = 0.0;
            break;
        }
        e = kv_div(ctx, len, ws.r0, r0norm, ws.V);
        if (e) return e;
        p[0] = r0norm;
        int nvecs = 0;
        for (int j = 0; j < m; ++j) {
            nvecs++;
            double *vj = ws.V + (size_t)j * len, *zj = ws.Z + (size_t)j * len;
            auto ta = clk::now();
            CDG_HIP_TRY(hipEventRecord(S.ev[depth][0], ctx->stream));
            e = prec_apply_dev(ctx, S.M, S.A, vj, zj);
            if (e) return e;
            if (need_inner) {
                // zr = v_j - A z_j ; inner solve ; z_j += dz   (:179-183)
                e = launch_mv(ctx, S.A, zj, ws.zr);
                if (!e) e = kv_sub(ctx, len, vj, ws.zr, ws.zr);
                int inner_it = 0;
                if (!e) e = fgmres_core(S, wss, depth + 1, ws.zr, ws.dz, inner->inner_nkrylov, inner->inner_tol,
                                        inner->inner_rtol, inner->inner_nmax, inner->inner_orthogonalization, nullptr,
                                        &inner_it);
                if (e) return e;
                e = kv_add(ctx, len, zj, ws.dz);
                if (e) return e;
            }
            CDG_HIP_TRY(hipEventRecord(S.ev[depth][1], ctx->stream));
            // w = A z_j, written into the slot of v_{j+1}   (:189)
            double *wj = ws.V + (size_t)(j + 1) * len;
            e = launch_mv(ctx, S.A, zj, wj);
            if (e) return e;
            CDG_HIP_TRY(hipEventRecord(S.ev[depth][2], ctx->stream));
            // orthogonalize; nb = ||w|| before it comes with the same fetch   (:191-210)
            double nrm2 = 0.0, normb2 = 0.0;
            int reorth = -1;
            e = orthogonalize(S, ws, orth, j + 1, wj, &H(0, j), &nrm2, &normb2, &reorth);
            if (e) return e;
            {   // the fetch synchronized the stream: the events are complete
                float ms_p = 0.f, ms_m = 0.f;
                CDG_HIP_TRY(hipEventElapsedTime(&ms_p, S.ev[depth][0], S.ev[depth][1]));
                CDG_HIP_TRY(hipEventElapsedTime(&ms_m, S.ev[depth][1], S.ev[depth][2]));
                const double tp = 1e-3 * ms_p, tm = 1e-3 * ms_m;
                S.t.precon += tp;
                S.t.mv += tm;
                S.t.dot += std::max(0.0, secs(ta, clk::now()) - tp - tm);
            }
            // selective re-orthogonalization, Giraud-Langou L_crit = 1   (:212-223);
            // CGS ran the test and the second pass on the device already
            double crit = 0.0;
            if (reorth < 0) {
                const double norm_before = std::sqrt(normb2);
                for (int i = 0; i <= j; ++i) crit += std::fabs(H(i, j));
                crit /= norm_before;
            } else if (reorth) {
                S.nreorth++;
            }
            if (reorth < 0 && crit >= 1.0) {
                auto tr = clk::now();
                e = orthogonalize(S, ws, orth, j + 1, wj, &H(0, j), &nrm2, nullptr);
                if (e) return e;
                S.nreorth++;
                S.t.dot += secs(tr, clk::now());
            }
            // h(j+1,j) = ||w|| ; v_{j+1} = w / h(j+1,j)   (:229-231)
            auto te = clk::now();
            H(j + 1, j) = std::sqrt(nrm2);
            e = kv_div_inplace(ctx, len, wj, H(j + 1, j));
            if (e) return e;
            S.t.norm += secs(te, clk::now());
            // previous Givens rotations, new rotation, rotate p   (:235-264)
            for (int i = 0; i < j; ++i) {
                const double hij = c[i] * H(i, j) + s[i] * H(i + 1, j);
                const double hipj = -s[i] * H(i, j) + c[i] * H(i + 1, j);
                H(i, j) = hij;
                H(i + 1, j) = hipj;
            }
            const double gam = std::hypot(H(j, j), H(j + 1, j));
            c[j] = H(j, j) / gam;
            s[j] = H(j + 1, j) / gam;
            H(j, j) = gam;
            H(j + 1, j) = 0.0;
            const double pj = c[j] * p[j], pjp = -s[j] * p[j];
            p[j] = pj;
            p[j + 1] = pjp;
            niter++;
            res = std::fabs(p[j + 1]);
            const bool ca = res < tol, cr = (res / r0norm_initial) < rtol, cn = (niter >= nmax) && (nmax > 0);
            if (!std::isfinite(res)) return cdg_fail(ctx, CDG_ERR_NAN, "non-finite FGMRES residual estimate");
            if (ca || cr || cn) { conv = true; break; }
        }
        // y = H(1:k,1:k)^-1 p(1:k) by back substitution of the triangular H   (:287-288)
        for (int i = nvecs - 1; i >= 0; --i) {
            double acc = p[i];
            for (int k = i + 1; k < nvecs; ++k) acc -= H(i, k) * y[k];
            y[i] = acc / H(i, i);
        }
        Workspace &w3 = ws;
        e = kv_combine(ctx, len, x, w3.x0, w3.Z, len, nvecs, y.data());
        if (e) return e;
        if (conv) break;
        e = kv_copy(ctx, len, x, w3.x0);
        if (e) return e;
        if (depth == 0) S.nrestart++;
    }
    if (niter_out) *niter_out = niter;
    return CDG_OK;
}

// ---------------------------------------------------------------------------
// k right-hand sides in lockstep: k independent runs of the outer FGMRES
// (type_fgmres.f90:81-310, inner FGMRES off), one per right-hand side, whose
// Arnoldi steps share launches: one k-vector ILU(0) apply and one k-vector
// matvec per step (each block read once for all k), the per-vector
// orthogonalization launches, and one device-to-host fetch of all vectors'
// scalars.  Each vector keeps its own Hessenberg matrix, rotations,
// residual estimate, restarts and stopping test, so it takes the iterations it
// takes alone; a vector that stops waits for the others' cycle to end.
// Vectors are laid out V[(j k + r) len], Z likewise.
struct MultiWs {
    int m = 0, k = 0;
    int64_t len = 0;
    double *V = nullptr, *Z = nullptr, *r0 = nullptr, *x0 = nullptr;
    void *y = nullptr;      // forward vectors of the k-vector apply
    void *base = nullptr;
    size_t bytes = 0;
};

int mws_alloc(cdg_ctx *ctx, MultiWs &ws, int m, int k, int64_t len) {
    ws.m = m;
    ws.k = k;
    ws.len = len;
    const size_t lv = (size_t)std::max<int64_t>(len, 1) * k;
    ws.bytes = (2 * (size_t)(m + 1) + 2) * lv * sizeof(double) + lv * sizeof(double);
    int e = dev_alloc(ctx, &ws.base, ws.bytes);
    if (e) return e;
    double *p = static_cast<double *>(ws.base);
    ws.V = p; p += (size_t)(m + 1) * lv;
    ws.Z = p; p += (size_t)(m + 1) * lv;
    ws.r0 = p; p += lv;
    ws.x0 = p; p += lv;
    ws.y = p;
    return CDG_OK;
}

// z_r = M^-1 v_r for r < k (vectors `len` apart): the k-vector row-slice
// apply for ILU(0), else one apply per vector.
int prec_apply_multi(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z, int64_t len, int k,
                     void *y) {
    if (M->type == CDG_PREC_IDENTITY) return kv_copy(ctx, len * k, v, z);
    if (!M->factored) return cdg_fail(ctx, CDG_ERR_STATE, "ILU0 apply before update");
    if (k > 1 && M->type == CDG_PREC_ILU0 && M->rs_ok) {
        int e = rs_prepare(ctx, M, A);
        if (e == CDG_ERR_UNSUPPORTED) {
            M->rs_ok = 0;
        } else {
            if (e) return e;
            M->napply += k;
            return launch_ilu_apply_rs_multi(ctx, M, v, z, y, len, k);
        }
    }
    for (int r = 0; r < k; ++r) {
        int e = prec_apply_dev(ctx, M, A, v + (size_t)r * len, z + (size_t)r * len);
        if (e) return e;
    }
    return CDG_OK;
}

struct Lane {
    std::vector<double> h, c, s, p, y;
    double r0norm_initial = 0.0, res = std::numeric_limits<double>::max();
    int niter = 0, nvecs = 0, nrestart = 0, nreorth = 0;
    bool done = false;      // converged (or stopped by nmax)
    bool in_cycle = false;  // still taking Arnoldi steps in this cycle
    bool conv = false;      // this cycle ended by the stopping test
    double &H(int m, int i, int j) { return h[(size_t)j * (m + 1) + i]; }
};

// Device ops of one vector's orthogonalization of w = V_jj against V_0..V_jj-1
// (V_i at stride vstride), results in d_h with orthogonalize()'s layout.
int orth_enqueue(cdg_ctx *ctx, int64_t len, int orth, int jj, double *w, const double *V, int64_t vstride,
                 double *d_h, bool with_normb, bool gated) {
    int e = CDG_OK;
    if (orth == CDG_MGS) {
        if (with_normb) e = kv_sumsq(ctx, len, w, d_h + jj);
        for (int i = 0; i < jj && !e; ++i) {
            e = kv_multidot(ctx, len, w, V + (size_t)i * vstride, vstride, 1, d_h + i);
            if (!e) e = kv_axpy_dev(ctx, len, w, V + (size_t)i * vstride, d_h + i);
        }
        if (!e) e = kv_sumsq(ctx, len, w, d_h + jj + 1);
        return e;
    }
    double *flag = gated ? d_h + jj + 2 : nullptr, *h2 = d_h + jj + 3;
    e = kv_multidot(ctx, len, w, V, vstride, with_normb ? jj + 1 : jj, d_h);
    if (!e) e = kv_cgs_update(ctx, len, w, V, vstride, jj, d_h, d_h + jj + 1, nullptr, flag);
    if (gated) {
        if (!e) e = kv_multidot(ctx, len, w, V, vstride, jj, h2, flag);
        if (!e) e = kv_cgs_update(ctx, len, w, V, vstride, jj, h2, h2 + jj, flag);
    }
    return e;
}

int fgmres_multi(Solver &S, MultiWs &ws, int k, const double *B, double *X, int m, double tol, double rtol, int nmax,
                 int orth, std::vector<Lane> &L) {
    cdg_ctx *ctx = S.ctx;
    const int64_t len = S.len;
    const int64_t kl = (int64_t)k * len;   // stride between basis steps
    if (m < 1) return cdg_fail(ctx, CDG_ERR_ARG, "nkrylov must be >= 1");
    const size_t hstride = 2 * (size_t)m + 8;
    int e = ctx_ensure_small(ctx, 16 + (size_t)m + hstride * k);
    if (e) return e;
    double *d_lane = ctx->d_small + 16 + m;   // beyond the scratch kv_combine uses
    L.assign(k, Lane());
    for (Lane &ln : L) {
        ln.h.assign((size_t)(m + 1) * m, 0.0);
        ln.c.assign(m + 1, 0.0);
        ln.s.assign(m + 1, 0.0);
        ln.p.assign(m + 1, 0.0);
        ln.y.assign(m + 1, 0.0);
    }
    e = kv_zero(ctx, kl, ws.x0);
    if (!e) e = kv_zero(ctx, kl, X);
    if (e) return e;
    auto any = [&](bool Lane::*f) { for (Lane &ln : L) if (ln.*f) return true; return false; };
    auto none_done = [&] { for (Lane &ln : L) if (!ln.done) return true; return false; };
    while (none_done()) {
        // r0 = b - A x0, v1 = r0 / ||r0||, p1 = ||r0|| for every vector still running   (:160-163)
        auto t0 = clk::now();
        e = launch_mv_multi(ctx, S.A, ws.x0, ws.r0, len, k);
        for (int r = 0; r < k && !e; ++r) {
            if (L[r].done) continue;
            e = kv_sub(ctx, len, B + (size_t)r * len, ws.r0 + (size_t)r * len, ws.r0 + (size_t)r * len);
            if (!e) e = kv_sumsq(ctx, len, ws.r0 + (size_t)r * len, d_lane + (size_t)r * hstride);
        }
        if (e) return e;
        CDG_HIP_TRY(hipMemcpyAsync(ctx->h_small, d_lane, sizeof(double) * hstride * k, hipMemcpyDeviceToHost,
                                   ctx->stream));
        e = ctx_wait(ctx);
        if (e) return e;
        S.t.norm += secs(t0, clk::now());
        std::vector<double> r0n(k, 0.0);
        for (int r = 0; r < k; ++r) {
            Lane &ln = L[r];
            ln.in_cycle = false;
            ln.conv = false;
            if (ln.done) continue;
            r0n[r] = std::sqrt(ctx->h_small[(size_t)r * hstride]);
            if (!std::isfinite(r0n[r])) return cdg_fail(ctx, CDG_ERR_NAN, "non-finite residual norm (vector %d)", r);
            if (ln.niter == 0) ln.r0norm_initial = r0n[r];
            std::fill(ln.h.begin(), ln.h.end(), 0.0);
            std::fill(ln.p.begin(), ln.p.end(), 0.0);
            std::fill(ln.c.begin(), ln.c.end(), 0.0);
            std::fill(ln.s.begin(), ln.s.end(), 0.0);
            if (r0n[r] == 0.0) {   // exact x0: x = x0 (the reference divides by zero here)
                e = kv_copy(ctx, len, ws.x0 + (size_t)r * len, X + (size_t)r * len);
                if (e) return e;
                ln.res = 0.0;
                ln.done = true;
                continue;
            }
            e = kv_div(ctx, len, ws.r0 + (size_t)r * len, r0n[r], ws.V + (size_t)r * len);
            if (e) return e;
            ln.p[0] = r0n[r];
            ln.nvecs = 0;
            ln.in_cycle = true;
        }
        for (int j = 0; j < m && any(&Lane::in_cycle); ++j) {
            double *Vj = ws.V + (size_t)j * kl, *Zj = ws.Z + (size_t)j * kl, *Wj = ws.V + (size_t)(j + 1) * kl;
            auto ta = clk::now();
            CDG_HIP_TRY(hipEventRecord(S.ev[0][0], ctx->stream));
            e = prec_apply_multi(ctx, S.M, S.A, Vj, Zj, len, k, ws.y);
            if (e) return e;
            CDG_HIP_TRY(hipEventRecord(S.ev[0][1], ctx->stream));
            e = launch_mv_multi(ctx, S.A, Zj, Wj, len, k);   // w = A z_j into the slot of v_{j+1}   (:189)
            if (e) return e;
            CDG_HIP_TRY(hipEventRecord(S.ev[0][2], ctx->stream));
            // orthogonalize every running vector, then one fetch   (:191-223)
            const bool gated = orth != CDG_MGS;
            for (int r = 0; r < k && !e; ++r)
                if (L[r].in_cycle)
                    e = orth_enqueue(ctx, len, orth, j + 1, Wj + (size_t)r * len, ws.V + (size_t)r * len, kl,
                                     d_lane + (size_t)r * hstride, true, gated);
            if (e) return e;
            CDG_HIP_TRY(hipMemcpyAsync(ctx->h_small, d_lane, sizeof(double) * hstride * k, hipMemcpyDeviceToHost,
                                       ctx->stream));
            e = ctx_wait(ctx);
            if (e) return e;
            {
                float ms_p = 0.f, ms_m = 0.f;
                CDG_HIP_TRY(hipEventElapsedTime(&ms_p, S.ev[0][0], S.ev[0][1]));
                CDG_HIP_TRY(hipEventElapsedTime(&ms_m, S.ev[0][1], S.ev[0][2]));
                S.t.precon += 1e-3 * ms_p;
                S.t.mv += 1e-3 * ms_m;
                S.t.dot += std::max(0.0, secs(ta, clk::now()) - 1e-3 * (ms_p + ms_m));
            }
            const int jj = j + 1;
            std::vector<int> redo;   // MGS vectors whose test asks for a second pass
            std::vector<double> nrm2(k, 0.0);
            for (int r = 0; r < k; ++r) {
                Lane &ln = L[r];
                if (!ln.in_cycle) continue;
                const double *hs = ctx->h_small + (size_t)r * hstride;
                for (int i = 0; i < jj; ++i) ln.H(m, i, j) += hs[i];
                const double normb2 = hs[jj];
                nrm2[r] = hs[jj + 1];
                if (gated) {
                    if (hs[jj + 2] != 0.0) {
                        for (int i = 0; i < jj; ++i) ln.H(m, i, j) += hs[jj + 3 + i];
                        nrm2[r] = hs[2 * jj + 3];
                        ln.nreorth++;
                    }
                } else {   // selective re-orthogonalization, Giraud-Langou L_crit = 1   (:212-223)
                    double crit = 0.0;
                    for (int i = 0; i <= j; ++i) crit += std::fabs(ln.H(m, i, j));
                    crit /= std::sqrt(normb2);
                    if (crit >= 1.0) redo.push_back(r);
                }
            }
            if (!redo.empty()) {
                for (int r : redo)
                    if (!e) e = orth_enqueue(ctx, len, orth, jj, Wj + (size_t)r * len, ws.V + (size_t)r * len, kl,
                                             d_lane + (size_t)r * hstride, false, false);
                if (e) return e;
                CDG_HIP_TRY(hipMemcpyAsync(ctx->h_small, d_lane, sizeof(double) * hstride * k,
                                           hipMemcpyDeviceToHost, ctx->stream));
                e = ctx_wait(ctx);
                if (e) return e;
                for (int r : redo) {
                    const double *hs = ctx->h_small + (size_t)r * hstride;
                    for (int i = 0; i < jj; ++i) L[r].H(m, i, j) += hs[i];
                    nrm2[r] = hs[jj + 1];
                    L[r].nreorth++;
                }
            }
            for (int r = 0; r < k; ++r) {
                Lane &ln = L[r];
                if (!ln.in_cycle) continue;
                ln.nvecs++;
                // h(j+1,j) = ||w|| ; v_{j+1} = w / h(j+1,j)   (:229-231)
                ln.H(m, j + 1, j) = std::sqrt(nrm2[r]);
                e = kv_div_inplace(ctx, len, Wj + (size_t)r * len, ln.H(m, j + 1, j));
                if (e) return e;
                // rotations   (:235-264)
                for (int i = 0; i < j; ++i) {
                    const double hij = ln.c[i] * ln.H(m, i, j) + ln.s[i] * ln.H(m, i + 1, j);
                    const double hipj = -ln.s[i] * ln.H(m, i, j) + ln.c[i] * ln.H(m, i + 1, j);
                    ln.H(m, i, j) = hij;
                    ln.H(m, i + 1, j) = hipj;
                }
                const double gam = std::hypot(ln.H(m, j, j), ln.H(m, j + 1, j));
                ln.c[j] = ln.H(m, j, j) / gam;
                ln.s[j] = ln.H(m, j + 1, j) / gam;
                ln.H(m, j, j) = gam;
                ln.H(m, j + 1, j) = 0.0;
                const double pj = ln.c[j] * ln.p[j], pjp = -ln.s[j] * ln.p[j];
                ln.p[j] = pj;
                ln.p[j + 1] = pjp;
                ln.niter++;
                ln.res = std::fabs(ln.p[j + 1]);
                if (!std::isfinite(ln.res))
                    return cdg_fail(ctx, CDG_ERR_NAN, "non-finite FGMRES residual estimate (vector %d)", r);
                const bool ca = ln.res < tol, cr = (ln.res / ln.r0norm_initial) < rtol,
                           cn = (ln.niter >= nmax) && (nmax > 0);
                if (ca || cr || cn) { ln.conv = true; ln.in_cycle = false; }
            }
        }
        // x = x0 + Z y per vector   (:287-294); converged vectors are done, the others restart
        for (int r = 0; r < k; ++r) {
            Lane &ln = L[r];
            if (ln.done || ln.nvecs == 0) continue;
            for (int i = ln.nvecs - 1; i >= 0; --i) {
                double acc = ln.p[i];
                for (int q = i + 1; q < ln.nvecs; ++q) acc -= ln.H(m, i, q) * ln.y[q];
                ln.y[i] = acc / ln.H(m, i, i);
            }
            e = kv_combine(ctx, len, X + (size_t)r * len, ws.x0 + (size_t)r * len, ws.Z + (size_t)r * len, kl,
                           ln.nvecs, ln.y.data());
            if (e) return e;
            if (ln.conv || !(ln.res > tol)) {
                ln.done = true;
            } else {
                e = kv_copy(ctx, len, X + (size_t)r * len, ws.x0 + (size_t)r * len);
                if (e) return e;
                ln.nrestart++;
            }
        }
    }
    return CDG_OK;
}

}  // namespace

extern "C" int cdg_fgmres_solve(cdg_ctx *ctx, const cdg_mat *A, cdg_prec *M, const double *b, double *x,
                                const cdg_fgmres_params *prm, cdg_fgmres_stats *st) {
    if (!ctx || !A || !M || !b || !x || !prm) return CDG_ERR_ARG;
    if (A->nrows != M->nrows || A->n != M->n) return cdg_fail(ctx, CDG_ERR_ARG, "matrix does not match preconditioner");
    cdg_fgmres_stats local;
    if (!st) st = &local;
    std::memset(st, 0, sizeof(*st));
    auto t_start = clk::now();
    int e = CDG_OK;
    // M%update(A,b) when the controller asks (type_fgmres.f90:118-122)
    if (prm->update_preconditioner || !M->factored || M->stamp != A->stamp) {
        auto tf = clk::now();
        e = cdg_prec_update(ctx, M, A, nullptr);
        if (e) return e;
        st->factored = 1;
        st->t_factor = secs(tf, clk::now());
    }
    const int64_t len = (int64_t)A->nrows * A->n;
    VecArg bi, xo;
    e = stage_in(ctx, 0, b, len, bi);
    if (!e) e = stage_in(ctx, 1, x, len, xo);
    if (!e) e = upload_in(ctx, bi);
    if (e) return e;
    Solver S{ctx, A, M, len};
    for (auto &lv : S.ev)
        for (hipEvent_t &ev : lv)
            if (!e && hipEventCreate(&ev) != hipSuccess) e = cdg_fail(ctx, CDG_ERR_HIP, "hipEventCreate failed");
    auto destroy_events = [&] {
        for (auto &lv : S.ev)
            for (hipEvent_t &ev : lv)
                if (ev) (void)hipEventDestroy(ev);
    };
    if (e) { destroy_events(); return e; }
    std::vector<Workspace> wss(2);   // outer and inner level
    int niter = 0;
    e = fgmres_core(S, wss, 0, bi.dev, xo.dev, prm->nkrylov, prm->tol, prm->rtol, prm->nmax, prm->orthogonalization,
                    prm->inner_fgmres ? prm : nullptr, &niter);
    // residual_norm = ||b - A x||   (type_fgmres.f90:305,515)
    if (!e) {
        e = launch_mv(ctx, A, xo.dev, wss[0].w);
        if (!e) e = kv_sub(ctx, len, bi.dev, wss[0].w, wss[0].w);
        if (!e) e = kv_sumsq(ctx, len, wss[0].w, ctx->d_small);
        if (!e) e = fetch_small(ctx, 1);
        if (!e) st->residual_norm = std::sqrt(ctx->h_small[0]);
    }
    if (!e) e = finish_out(ctx, xo);
    for (Workspace &w : wss) ws_free(ctx, w);
    destroy_events();
    st->niter = niter;
    st->nrestart = S.nrestart;
    st->nreorth = S.nreorth;
    st->t_precon = S.t.precon;
    st->t_mv = S.t.mv;
    st->t_dot = S.t.dot;
    st->t_norm = S.t.norm;
    st->t_total = secs(t_start, clk::now());
    if (!e && !std::isfinite(st->residual_norm)) e = cdg_fail(ctx, CDG_ERR_NAN, "non-finite final residual");
    if (e == CDG_ERR_NAN) {
        const int e2 = df_check(ctx, M);
        if (e2) return e2;
    }
    return e;
}

extern "C" int cdg_fgmres_solve_multi(cdg_ctx *ctx, const cdg_mat *A, cdg_prec *M, int32_t k, const double *b, double *x,
                                      const cdg_fgmres_params *prm, cdg_fgmres_stats *st) {
    if (!ctx || !A || !M || !b || !x || !prm || k < 1) return CDG_ERR_ARG;
    if (A->nrows != M->nrows || A->n != M->n) return cdg_fail(ctx, CDG_ERR_ARG, "matrix does not match preconditioner");
    const int64_t len = (int64_t)A->nrows * A->n;
    if (k == 1 || prm->inner_fgmres) {   // one vector, or nested FGMRES: one solve per vector
        for (int r = 0; r < k; ++r) {
            cdg_fgmres_params p1 = *prm;
            if (r > 0) p1.update_preconditioner = 0;   // the first solve's factor serves the others
            int e = cdg_fgmres_solve(ctx, A, M, b + (size_t)r * len, x + (size_t)r * len, &p1, st ? st + r : nullptr);
            if (e) return e;
        }
        return CDG_OK;
    }
    std::vector<cdg_fgmres_stats> local(k);
    auto t_start = clk::now();
    int e = CDG_OK;
    double t_factor = 0.0;
    int factored = 0;
    if (prm->update_preconditioner || !M->factored || M->stamp != A->stamp) {
        auto tf = clk::now();
        e = cdg_prec_update(ctx, M, A, nullptr);
        if (e) return e;
        factored = 1;
        t_factor = secs(tf, clk::now());
    }
    VecArg bi, xo;
    e = stage_in(ctx, 0, b, len * k, bi);
    if (!e) e = stage_in(ctx, 1, x, len * k, xo);
    if (!e) e = upload_in(ctx, bi);
    if (e) return e;
    Solver S{ctx, A, M, len};
    for (auto &lv : S.ev)
        for (hipEvent_t &ev : lv)
            if (!e && hipEventCreate(&ev) != hipSuccess) e = cdg_fail(ctx, CDG_ERR_HIP, "hipEventCreate failed");
    MultiWs ws;
    if (!e) e = mws_alloc(ctx, ws, prm->nkrylov, k, len);
    std::vector<Lane> lanes;
    if (!e) e = fgmres_multi(S, ws, k, bi.dev, xo.dev, prm->nkrylov, prm->tol, prm->rtol, prm->nmax,
                             prm->orthogonalization, lanes);
    // residual_norm = ||b - A x|| per vector   (type_fgmres.f90:305,515)
    if (!e) {
        const size_t hstride = 2 * (size_t)prm->nkrylov + 8;
        double *d_lane = ctx->d_small + 16 + prm->nkrylov;
        e = launch_mv_multi(ctx, A, xo.dev, ws.r0, len, k);
        for (int r = 0; r < k && !e; ++r) {
            e = kv_sub(ctx, len, bi.dev + (size_t)r * len, ws.r0 + (size_t)r * len, ws.r0 + (size_t)r * len);
            if (!e) e = kv_sumsq(ctx, len, ws.r0 + (size_t)r * len, d_lane + (size_t)r * hstride);
        }
        if (!e) {
            hipError_t he = hipMemcpyAsync(ctx->h_small, d_lane, sizeof(double) * hstride * k, hipMemcpyDeviceToHost,
                                           ctx->stream);
            e = he == hipSuccess ? ctx_wait(ctx) : cdg_fail(ctx, CDG_ERR_HIP, "%s", hipGetErrorString(he));
        }
        for (int r = 0; r < k && !e; ++r) local[r].residual_norm = std::sqrt(ctx->h_small[(size_t)r * hstride]);
    }
    if (!e) e = finish_out(ctx, xo);
    dev_free(ctx, ws.base, ws.bytes);
    for (auto &lv : S.ev)
        for (hipEvent_t &ev : lv)
            if (ev) (void)hipEventDestroy(ev);
    const double t_total = secs(t_start, clk::now());
    for (int r = 0; r < k && r < (int)lanes.size(); ++r) {
        cdg_fgmres_stats &o = local[r];
        o.niter = lanes[r].niter;
        o.nrestart = lanes[r].nrestart;
        o.nreorth = lanes[r].nreorth;
        o.factored = factored;
        o.t_factor = t_factor;
        o.t_precon = S.t.precon;
        o.t_mv = S.t.mv;
        o.t_dot = S.t.dot;
        o.t_norm = S.t.norm;
        o.t_total = t_total;
        if (!e && !std::isfinite(o.residual_norm)) e = cdg_fail(ctx, CDG_ERR_NAN, "non-finite final residual");
    }
    if (st) std::memcpy(st, local.data(), sizeof(cdg_fgmres_stats) * k);
    return e;
}

extern "C" int cdg_mv_multi(cdg_ctx *ctx, const cdg_mat *A, int32_t k, const double *X, double *Y) {
    if (!ctx || !A || !X || !Y || k < 1) return CDG_ERR_ARG;
    const int64_t len = (int64_t)A->nrows * A->n;
    VecArg xi, yo;
    int e = stage_in(ctx, 0, X, len * k, xi);
    if (!e) e = stage_in(ctx, 1, Y, len * k, yo);
    if (!e) e = upload_in(ctx, xi);
    if (!e) e = launch_mv_multi(ctx, A, xi.dev, yo.dev, len, k);
    if (!e) e = finish_out(ctx, yo);
    return e;
}

extern "C" int cdg_prec_apply_multi(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, int32_t k, const double *V, double *Z) {
    if (!ctx || !M || !A || !V || !Z || k < 1) return CDG_ERR_ARG;
    if (A->nrows != M->nrows || A->n != M->n) return cdg_fail(ctx, CDG_ERR_ARG, "matrix does not match preconditioner");
    const int64_t len = (int64_t)A->nrows * A->n;
    const size_t ybytes = (M->precision == CDG_FP64 ? 8 : 4) * (size_t)len * k;
    if (M->type == CDG_PREC_ILU0 && M->ym_bytes < ybytes) {   // forward vectors of the k-vector apply
        if (M->ym) dev_free(ctx, M->ym, M->ym_bytes);
        M->ym = nullptr;
        M->ym_bytes = 0;
        int e = dev_alloc(ctx, &M->ym, ybytes);
        if (e) return e;
        M->ym_bytes = ybytes;
    }
    VecArg vi, zo;
    int e = stage_in(ctx, 0, V, len * k, vi);
    if (!e) e = stage_in(ctx, 1, Z, len * k, zo);
    if (!e) e = upload_in(ctx, vi);
    if (!e) e = prec_apply_multi(ctx, M, A, vi.dev, zo.dev, len, k, M->ym);
    if (!e) e = finish_out(ctx, zo);
    return e;
}

// ---- device-resident vectors (chidg_vector_t operations, type_chidg_vector.f90) ----
extern "C" int cdg_vec_create(cdg_ctx *ctx, int64_t len, double **out) {
    if (!ctx || !out || len < 0) return CDG_ERR_ARG;
    *out = nullptr;
    CDG_HIP_TRY(hipSetDevice(ctx->device));
    void *p = nullptr;
    int e = dev_alloc(ctx, &p, sizeof(double) * (size_t)std::max<int64_t>(len, 1));
    if (!e) e = kv_zero(ctx, len, static_cast<double *>(p));
    if (!e) e = ctx_wait(ctx);
    if (e) { if (p) dev_free(ctx, p, 0); return e; }
    *out = static_cast<double *>(p);
    return CDG_OK;
}

extern "C" int cdg_vec_destroy(cdg_ctx *ctx, double *v) {
    if (!ctx) return CDG_ERR_ARG;
    if (v) {
        CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
        dev_free(ctx, v, 0);
    }
    return CDG_OK;
}

extern "C" int cdg_vec_copy(cdg_ctx *ctx, int64_t len, const double *src, double *dst) {
    if (!ctx || !src || !dst || len < 0) return CDG_ERR_ARG;
    CDG_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)len, hipMemcpyDefault, ctx->stream));
    return ctx_wait(ctx);
}

extern "C" int cdg_vec_axpby(cdg_ctx *ctx, int64_t len, double a, const double *x, double b, double *y) {
    if (!ctx || !x || !y || len < 0) return CDG_ERR_ARG;
    if (!is_device_ptr(x) || !is_device_ptr(y)) return cdg_fail(ctx, CDG_ERR_ARG, "cdg_vec_axpby needs device pointers");
    int e = kv_axpby(ctx, len, a, x, b, y);
    if (!e) e = ctx_wait(ctx);
    return e;
}

extern "C" int cdg_vec_dot(cdg_ctx *ctx, int64_t len, const double *x, const double *y, double *out) {
    if (!ctx || !x || !y || !out || len < 0) return CDG_ERR_ARG;
    VecArg xi, yi;
    int e = stage_in(ctx, 0, x, len, xi);
    if (!e) e = stage_in(ctx, 1, y, len, yi);
    if (!e) e = upload_in(ctx, xi);
    if (!e) e = upload_in(ctx, yi);
    if (!e) e = kv_multidot(ctx, len, xi.dev, yi.dev, len, 1, ctx->d_small);
    if (!e) e = fetch_small(ctx, 1);
    if (!e) *out = ctx->h_small[0];
    return e;
}

extern "C" int cdg_vec_norm(cdg_ctx *ctx, int64_t len, const double *x, double *out) {
    if (!ctx || !x || !out || len < 0) return CDG_ERR_ARG;
    VecArg xi;
    int e = stage_in(ctx, 0, x, len, xi);
    if (!e) e = upload_in(ctx, xi);
    if (!e) e = kv_sumsq(ctx, len, xi.dev, ctx->d_small);
    if (!e) e = fetch_small(ctx, 1);
    if (!e) *out = std::sqrt(ctx->h_small[0]);
    return e;
}

// ---- DG Euler residual on the device (mod_spatial.f90:40-197) ----------------
extern "C" int cdg_euler_residual_gpu(cdg_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, int32_t order, const double *q,
                                      double *R) {
    if (!ctx || !q || !R) return CDG_ERR_ARG;
    if (nx < 2 || nx > 256 || ny != nx || nz != nx || order < 0 || order > 3) return CDG_ERR_ARG;
    const int64_t len = (int64_t)nx * ny * nz * 5 * (order + 1) * (order + 1) * (order + 1);
    VecArg qi, ro;
    int e = stage_in(ctx, 0, q, len, qi);
    if (!e) e = stage_in(ctx, 1, R, len, ro);
    if (!e) e = upload_in(ctx, qi);
    if (!e) e = launch_euler_residual(ctx, nx, ny, nz, order, qi.dev, ro.dev);
    if (!e) e = finish_out(ctx, ro);
    if (!e) {   // a state with negative pressure gives a non-finite residual, as on the host
        e = kv_sumsq(ctx, len, ro.dev, ctx->d_small);
        if (!e) e = fetch_small(ctx, 1);
        if (!e && !std::isfinite(ctx->h_small[0])) e = cdg_fail(ctx, CDG_ERR_NAN, "non-finite Euler residual");
    }
    return e;
}
