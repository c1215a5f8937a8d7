// Block-row streaming kernels for gfx950: the block-sparse matvec and the
// ILU(0) forward/backward sweeps.
//
// Both stream dense n x n column-major blocks from HBM.  A workgroup owns one
// block row.  Thread t reads V consecutive elements (16 bytes: 2 FP64, 4 FP32,
// 8 BF16) at flat offsets V*t + V*T*m of every block; T is chosen so that V*T
// is a multiple of n, hence each thread always meets the same V matrix rows
// and keeps V accumulators in registers while its column index advances by
// V*T/n per step.  Loads are 16-byte, fully coalesced; the parent vectors sit
// in LDS; partial sums are reduced through LDS in a fixed order
// (deterministic results).
#include "cdg_internal.h"

#include <algorithm>

namespace {

constexpr int kMaxThreads = 1024;

typedef double nat_f64x2 __attribute__((ext_vector_type(2)));
typedef float nat_f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t nat_u32x4 __attribute__((ext_vector_type(4)));

// ---- 16-byte vector loads converted to the accumulation type ----------------
template <typename S, typename Acc> struct Ld16;

template <> struct Ld16<double, double> {
    static constexpr int V = 2;
    __device__ static __forceinline__ void load(const double *p, double (&o)[2]) {
        const nat_f64x2 t = __builtin_nontemporal_load(reinterpret_cast<const nat_f64x2 *>(p));
        o[0] = t.x; o[1] = t.y;
    }
};
template <> struct Ld16<float, float> {
    static constexpr int V = 4;
    __device__ static __forceinline__ void load(const float *p, float (&o)[4]) {
        const nat_f32x4 t = __builtin_nontemporal_load(reinterpret_cast<const nat_f32x4 *>(p));
        o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
    }
};
template <> struct Ld16<cdg_bf16, float> {
    static constexpr int V = 8;
    __device__ static __forceinline__ void load(const cdg_bf16 *p, float (&o)[8]) {
        const nat_u32x4 t = __builtin_nontemporal_load(reinterpret_cast<const nat_u32x4 *>(p));
        const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            o[2 * i] = __uint_as_float(w[i] << 16);
            o[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
        }
    }
};

// Per-thread view of the flat element walk.
template <int V> struct Walk {
    int e0;        // first flat element of this thread
    int c0[V];     // initial column of each of the V elements
    int dc;        // column advance per step
    int step;      // V*T
};

template <int V> __device__ __forceinline__ Walk<V> make_walk(int tid, int T, int n, int dc) {
    Walk<V> w;
    w.e0 = V * tid;
    w.dc = dc;
    w.step = V * T;
#pragma unroll
    for (int v = 0; v < V; ++v) w.c0[v] = (w.e0 + v) / n;
    return w;
}

// acc[v] += sum over this thread's elements of blk(e) * xs(col(e)).
template <typename S, typename Acc>
__device__ __forceinline__ void stream_block(const S *__restrict__ blk, const Acc *__restrict__ xs, int nn,
                                             const Walk<Ld16<S, Acc>::V> &w, Acc *acc) {
    constexpr int V = Ld16<S, Acc>::V;
    const int mfull = nn / w.step;           // steps where every element is in range
    int m = 0;
    for (; m + 4 <= mfull; m += 4) {
        Acc a[4][V];
#pragma unroll
        for (int u = 0; u < 4; ++u) Ld16<S, Acc>::load(blk + w.e0 + (m + u) * w.step, a[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] += a[u][v] * xs[w.c0[v] + (m + u) * w.dc];
    }
    for (; m < mfull; ++m) {
        Acc a[V];
        Ld16<S, Acc>::load(blk + w.e0 + m * w.step, a);
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] += a[v] * xs[w.c0[v] + m * w.dc];
    }
    const int e = w.e0 + mfull * w.step;
    if (e < nn) {   // tail: the 16-byte load stays inside the padded slot
        Acc a[V];
        Ld16<S, Acc>::load(blk + e, a);
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (e + v < nn) acc[v] += a[v] * xs[w.c0[v] + mfull * w.dc];
    }
}

// Writes acc to red[V*t+v]; after a barrier, row r's sum is red[r + n*q], q < V*T/n.
template <typename Acc, int V>
__device__ __forceinline__ void park(Acc *red, int tid, const Acc *acc) {
#pragma unroll
    for (int v = 0; v < V; ++v) red[V * tid + v] = acc[v];
}

template <typename Acc>
__device__ __forceinline__ Acc row_sum(const Acc *red, int r, int n, int parts) {
    Acc s = 0;
    for (int q = 0; q < parts; ++q) s += red[r + n * q];
    return s;
}

// ============================================================================
// K1: y_i = sum_b A_b x_col(b)   (chidg_mv, operator_chidg_mv.f90:200-374)
// ============================================================================
__global__ void __launch_bounds__(kMaxThreads)
k_mv(const double *__restrict__ blocks, int64_t ld, const int32_t *__restrict__ row_ptr,
     const int32_t *__restrict__ col, const double *__restrict__ x, double *__restrict__ y, int n,
     int dc, int chunk) {
    extern __shared__ double smem[];
    constexpr int V = 2;
    const int T = blockDim.x, tid = threadIdx.x, row = blockIdx.x;
    const int nn = n * n;
    double *xs = smem;                       // chunk * n
    double *red = smem + (size_t)chunk * n;  // V*T
    const Walk<V> w = make_walk<V>(tid, T, n, dc);
    double acc[V] = {0.0, 0.0};
    const int b0 = row_ptr[row], b1 = row_ptr[row + 1];
    for (int cb = b0; cb < b1; cb += chunk) {
        const int nb = min(chunk, b1 - cb);
        if (cb != b0) __syncthreads();
        for (int idx = tid; idx < nb * n; idx += T) {
            const int k = idx / n, c = idx - k * n;
            xs[idx] = x[(size_t)col[cb + k] * n + c];
        }
        __syncthreads();
        for (int k = 0; k < nb; ++k)
            stream_block<double, double>(blocks + (size_t)(cb + k) * ld, xs + k * n, nn, w, acc);
    }
    park<double, V>(red, tid, acc);
    __syncthreads();
    const int parts = V * T / n;
    for (int r = tid; r < n; r += T) y[(size_t)row * n + r] = row_sum(red, r, n, parts);
}

// ============================================================================
// K3: ILU(0) apply as one persistent dataflow kernel (precon_ILU0.f90:286-411)
//   task t <  nrows : forward row  tasks[t]:  z_i = v_i - sum_j L_ij z_j
//   task t >= nrows : backward row tasks[t]:  z_i = Dinv_i (z_i - sum_k U_ik z_k)
// Tasks are taken in topological order from an atomic counter; a task waits
// only on rows of earlier tasks (fdone/bdone flags == epoch), so the sweep has
// no level barriers and needs no co-residency guarantee.  Every wait is
// bounded: on timeout the kernel records the fact and drains.
// ============================================================================
struct SweepArgs {
    const void *L;           // nlo slots
    const void *U;           // upper slots (A's blocks for FP64)
    const void *Dinv;        // nrows slots
    int64_t ld, ldu;
    const int32_t *lo_ptr, *lo_col;          // L slot = lower index q
    const int32_t *up_ptr, *up_idx, *up_col; // U slot = up_idx[u]
    const int32_t *tasks;
    int32_t *fdone, *bdone, *counters;
    int32_t epoch;
    const double *v;
    double *z;
    int n, nrows, dc, maxdep;
};

constexpr int kSpinLimit = 1 << 21;   // polls (each followed by s_sleep) before giving up

// Waits for *f == epoch.  Gives up after kSpinLimit polls, or at once when
// another wait of this launch has already given up (*abort != 0), so a broken
// dependency drains the grid instead of hanging it.
__device__ __forceinline__ bool wait_flag(int32_t *f, int32_t epoch, int32_t *abort) {
    for (int it = 0; it < kSpinLimit; ++it) {
        if (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) return true;
        if ((it & 255) == 255 && __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        __builtin_amdgcn_s_sleep(2);
    }
    return false;
}

// Next task index, or -1 when the tickets are exhausted or the schedule
// entry is out of range (flagged in the status words).
__device__ __forceinline__ int take_ticket(const SweepArgs &a) {
    const int t = atomicAdd(&a.counters[0], 1);
    if (t < 0 || t >= 2 * a.nrows) return -1;
    const int r = a.tasks[t];
    if (r < 0 || r >= a.nrows) {
        atomicOr(&a.counters[2], 2);
        atomicOr(&a.counters[3], 2);
        return -1;
    }
    return t;
}

template <typename S, typename Acc>
__global__ void __launch_bounds__(kMaxThreads) k_ilu_sweep(SweepArgs a) {
    constexpr int V = Ld16<S, Acc>::V;
    extern __shared__ unsigned char smem_raw[];
    Acc *xs = reinterpret_cast<Acc *>(smem_raw);           // maxdep * n
    Acc *red = xs + (size_t)a.maxdep * a.n;                 // V*T
    Acc *tv = red + (size_t)V * blockDim.x;                 // n
    __shared__ int32_t task_sh;
    const int T = blockDim.x, tid = threadIdx.x, n = a.n, nn = n * n;
    const Walk<V> w = make_walk<V>(tid, T, n, a.dc);
    const int parts = V * T / n;
    const S *Lp = static_cast<const S *>(a.L);
    const S *Up = static_cast<const S *>(a.U);
    const S *Dp = static_cast<const S *>(a.Dinv);
    // Loop shape matters: the tid-0 work (publish the finished row, take the
    // next ticket) sits in one block followed by an unconditional barrier, so
    // the back-edge is convergent.  With tid-0 blocks at both the end and the
    // start of the body, the structurizer merges them across the back-edge and
    // the other lanes re-enter the next barrier before tid 0 has published the
    // row and taken a ticket: the workgroup never drains.
    if (tid == 0) task_sh = take_ticket(a);
    __syncthreads();
    for (;;) {
        const int task = __builtin_amdgcn_readfirstlane(task_sh);
        if (task < 0) break;   // tickets exhausted (or a corrupt schedule: flagged, drained)
        const bool fwd = task < a.nrows;
        const int row = __builtin_amdgcn_readfirstlane(a.tasks[task]);
        const int d0 = fwd ? a.lo_ptr[row] : a.up_ptr[row];
        const int d1 = fwd ? a.lo_ptr[row + 1] : a.up_ptr[row + 1];
        const int32_t *dcol = fwd ? a.lo_col : a.up_col;
        if (tid == 0) {
            bool ok = true;
            if (!fwd) ok = wait_flag(&a.fdone[row], a.epoch, &a.counters[2]);
            for (int d = d0; d < d1 && ok; ++d)
                ok = wait_flag(fwd ? &a.fdone[dcol[d]] : &a.bdone[dcol[d]], a.epoch, &a.counters[2]);
            if (!ok) { atomicOr(&a.counters[2], 1); atomicOr(&a.counters[3], 1); }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        const int nd = d1 - d0;
        for (int idx = tid; idx < nd * n; idx += T) {
            const int k = idx / n, c = idx - k * n;
            xs[idx] = static_cast<Acc>(a.z[(size_t)dcol[d0 + k] * n + c]);
        }
        __syncthreads();
        Acc acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = 0;
        for (int k = 0; k < nd; ++k) {
            const S *blk = fwd ? Lp + (size_t)(d0 + k) * a.ld : Up + (size_t)a.up_idx[d0 + k] * a.ldu;
            stream_block<S, Acc>(blk, xs + k * n, nn, w, acc);
        }
        park<Acc, V>(red, tid, acc);
        __syncthreads();
        double *zi = a.z + (size_t)row * n;
        if (fwd) {
            const double *vi = a.v + (size_t)row * n;
            for (int r = tid; r < n; r += T) zi[r] = vi[r] - static_cast<double>(row_sum(red, r, n, parts));
        } else {
            for (int r = tid; r < n; r += T) tv[r] = static_cast<Acc>(zi[r]) - row_sum(red, r, n, parts);
            __syncthreads();
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] = 0;
            stream_block<S, Acc>(Dp + (size_t)row * a.ld, tv, nn, w, acc);
            park<Acc, V>(red, tid, acc);
            __syncthreads();
            for (int r = tid; r < n; r += T) zi[r] = static_cast<double>(row_sum(red, r, n, parts));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) {
            __hip_atomic_store(fwd ? &a.fdone[row] : &a.bdone[row], a.epoch, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
            task_sh = take_ticket(a);
        }
        __syncthreads();
    }
}

// Arms the sweep: ticket counter and this launch's abort flag to zero (a
// kernel rather than a memset so the pair can be captured in a graph).
// counters[3] is the sticky status the host reads and clears.
__global__ void k_sweep_arm(int32_t *counters) {
    if (threadIdx.x < 3) counters[threadIdx.x] = 0;
}

// Identity preconditioner / widening copy: dst[k] = (double) src[k]
template <typename S>
__global__ void k_widen(const S *__restrict__ src, double *__restrict__ dst, int64_t count) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x)
        dst[k] = static_cast<double>(src[k]);
}
template <>
__global__ void k_widen<cdg_bf16>(const cdg_bf16 *__restrict__ src, double *__restrict__ dst, int64_t count) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x)
        dst[k] = static_cast<double>(__uint_as_float(static_cast<uint32_t>(src[k]) << 16));
}

}  // namespace

// ---- launch shape -------------------------------------------------------------
StreamShape stream_shape(int n, int V) {
    int g = n;
    {   // g = n / gcd(n, V)
        int a = n, b = V;
        while (b) { int t = a % b; a = b; b = t; }
        g = n / a;
    }
    StreamShape best;
    double best_score = -1.0;
    for (int T = g; T <= kMaxThreads; T += g) {
        if (T < 64 && T + g <= kMaxThreads) continue;
        const int waves = (T + 63) / 64;
        const double util = (double)T / (64.0 * waves);
        const double score = util - 1e-4 * std::abs(T - 640) / 64.0;
        if (score > best_score) { best_score = score; best.T = T; }
    }
    if (best.T == 0) best.T = g <= kMaxThreads ? g : 0;
    best.dc = best.T ? V * best.T / n : 0;
    return best;
}

int launch_mv(cdg_ctx *ctx, const cdg_mat *A, const double *x, double *y) {
    if (A->nrows == 0) return CDG_OK;
    const StreamShape &s = A->shape64;
    if (s.T == 0) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "block size %d not supported", A->n);
    const int V = 2;
    int chunk = std::max(1, std::min<int>(A->max_row_blocks, (32 * 1024) / (8 * A->n)));
    size_t lds = sizeof(double) * ((size_t)chunk * A->n + (size_t)V * s.T);
    hipLaunchKernelGGL(k_mv, dim3(A->nrows), dim3(s.T), lds, ctx->stream, A->blocks, A->ld, A->d_row_ptr,
                       A->d_col, x, y, A->n, s.dc, chunk);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int launch_ilu_apply(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    if (M->nrows == 0) return CDG_OK;
    SweepArgs a;
    a.L = M->L;
    a.Dinv = M->Dinv;
    a.ld = M->ld;
    if (M->precision == CDG_FP64) { a.U = A->blocks; a.ldu = A->ld; }
    else { a.U = M->U; a.ldu = M->ld; }
    a.lo_ptr = M->d_lo_ptr; a.lo_col = M->d_lo_col;
    a.up_ptr = M->d_up_ptr; a.up_idx = M->d_up_idx; a.up_col = M->d_up_col;
    a.tasks = M->d_tasks;
    a.fdone = M->d_flags; a.bdone = M->d_flags + M->nrows;
    a.counters = M->d_counters;
    a.epoch = ++M->epoch;
    a.v = v; a.z = z;
    a.n = M->n; a.nrows = M->nrows; a.dc = M->shape.dc;
    a.maxdep = M->maxdep;
    const int T = M->shape.T;
    size_t accsz = M->precision == CDG_FP64 ? sizeof(double) : sizeof(float);
    int V = M->precision == CDG_FP64 ? 2 : (M->precision == CDG_FP32 ? 4 : 8);
    size_t lds = accsz * ((size_t)a.maxdep * M->n + (size_t)V * T + (size_t)M->n);
    if (lds > 64 * 1024) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "ILU sweep needs %zu bytes of LDS", lds);
    dim3 grid(M->sweep_grid), block(T);
    hipLaunchKernelGGL(k_sweep_arm, dim3(1), dim3(64), 0, ctx->stream, M->d_counters);
    if (M->precision == CDG_FP64)
        hipLaunchKernelGGL((k_ilu_sweep<double, double>), grid, block, lds, ctx->stream, a);
    else if (M->precision == CDG_FP32)
        hipLaunchKernelGGL((k_ilu_sweep<float, float>), grid, block, lds, ctx->stream, a);
    else
        hipLaunchKernelGGL((k_ilu_sweep<cdg_bf16, float>), grid, block, lds, ctx->stream, a);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int launch_widen(cdg_ctx *ctx, const void *src, int precision, double *dst, int64_t count) {
    if (count == 0) return CDG_OK;
    dim3 grid((unsigned)std::min<int64_t>((count + 255) / 256, 4096)), block(256);
    if (precision == CDG_FP64)
        hipLaunchKernelGGL(k_widen<double>, grid, block, 0, ctx->stream, static_cast<const double *>(src), dst, count);
    else if (precision == CDG_FP32)
        hipLaunchKernelGGL(k_widen<float>, grid, block, 0, ctx->stream, static_cast<const float *>(src), dst, count);
    else
        hipLaunchKernelGGL(k_widen<cdg_bf16>, grid, block, 0, ctx->stream, static_cast<const cdg_bf16 *>(src), dst, count);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}
