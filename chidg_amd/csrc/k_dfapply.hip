// ILU(0) apply as one persistent dataflow launch (precon_ILU0.f90:286-411).
//
// The sweep is cut into tasks (stage, row, column slice):
//   F  (forward)          P_i[s]  = sum_j L_ij[:, s] z_j[s]     z_i = v_i - sum_s P_i[s]
//   BA (backward, part 1) QA_i[s] = sum_k U_ik[:, s] y_k[s]     t_i = z_i - sum_s QA_i[s]
//   BB (backward, part 2) PB_i[s] = Dinv_i[:, s] t_i[s]         y_i = sum_s PB_i[s]
// where [:, s] is the slice s of a block's columns (one contiguous range of a
// column-major block) and x[s] the matching entries of a vector.  Each task
// writes one partial n-vector; a row's value is the fixed-order sum of its
// partials, rebuilt by every consumer, so results do not depend on timing.
//
// Tasks are numbered in a topological order (forward levels ascending, then
// per backward level its BA tasks followed by its BB tasks) and handed out by
// a ticket counter.  A workgroup that holds ticket t first issues the loads of
// its block slices into registers, then waits until the rows it reads are
// complete (per-(stage,row) counters that reach epoch*S), then finishes the
// product.  A task's inputs always carry smaller tickets, all of which were
// handed to running workgroups, so the lowest unfinished ticket can always
// proceed and the launch drains whatever the residency.  Every wait is
// bounded: past the limit the workgroup raises the error word and leaves,
// and the others leave at their next check.
//
// Cross-workgroup data (partials, counters) moves with agent-scope atomic
// loads and stores, which are coherent across the XCDs' L2 caches; a
// producer waits for its partial's stores to complete before it increments
// the row's counter.
//
// Measured on 16^3 P2 this schedule is not faster than the level schedule:
// a dependency hop costs 3-5 us either way (one launch, or one
// store/counter/poll/load chain), and the sweep has 138 hops.
#include "cdg_internal.h"

#include <algorithm>

namespace {

typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16-byte raw vector of each storage type and its expansion to V accumulators
template <typename S> struct Raw;
template <> struct Raw<double> {
    typedef f64x2 T;
    static constexpr int V = 2;
    __device__ static __forceinline__ void get(const T &r, double (&o)[2]) { o[0] = r.x; o[1] = r.y; }
};
template <> struct Raw<float> {
    typedef f32x4 T;
    static constexpr int V = 4;
    __device__ static __forceinline__ void get(const T &r, float (&o)[4]) { o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = r.w; }
};
template <> struct Raw<cdg_bf16> {
    typedef u32x4 T;
    static constexpr int V = 8;
    __device__ static __forceinline__ void get(const T &r, float (&o)[8]) {
        const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            o[2 * i] = __uint_as_float(w[i] << 16);
            o[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
        }
    }
};

__device__ __forceinline__ double ld_co(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_co(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_co(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

enum { ST_F = 0, ST_BA = 1, ST_BB = 2 };


struct DfArgs {
    const void *L, *U, *Dinv;
    int64_t ldl, ldu;                     // slot strides (elements): L and Dinv, U
    const int32_t *lo_ptr, *lo_col;       // lower entries of a row (L slot = entry index)
    const int32_t *up_ptr, *up_idx, *up_col;   // upper entries (U slot = up_idx)
    const int32_t *tasks;                 // ((row*S + slice) << 2) | stage, in ticket order
    int32_t ntasks;
    int32_t *cnt;                         // [3][nrows] finished slices, cumulative over calls
    unsigned long long *ticket;
    unsigned long long calls;             // earlier calls since the counters were zeroed
    int32_t *err;
    const double *v;                      // the vector preconditioned
    double *part;                         // [3][nrows][S][n] partials
    int64_t plen;                         // nrows*S*n
    int32_t nrows, n, S, w, rounds, dc, target, maxdep, spin_limit;
};

// Value of the vector read by stage `st` at (row, c): the fixed-order sum of
// the producing stage's partials (plus the base for forward values).
__device__ __forceinline__ double df_view(const DfArgs &a, int st, int row, int c) {
    const size_t o = (size_t)row * a.S * a.n + c;
    if (st == ST_BA) {              // y_k = sum PB_k
        const double *pb = a.part + 2 * a.plen + o;
        double t = 0.0;
        for (int s = 0; s < a.S; ++s) t += ld_co(pb + (size_t)s * a.n);
        return t;
    }
    const double *p = a.part + o;   // z_j = v_j - sum P_j
    double t = 0.0;
    for (int s = 0; s < a.S; ++s) t += ld_co(p + (size_t)s * a.n);
    double x = a.v[(size_t)row * a.n + c] - t;
    if (st == ST_BB) {              // t_i = z_i - sum QA_i
        const double *q = a.part + a.plen + o;
        double u = 0.0;
        for (int s = 0; s < a.S; ++s) u += ld_co(q + (size_t)s * a.n);
        x -= u;
    }
    return x;
}

// At most 320 threads (5 waves); registers for two waves per SIMD (measured:
// fewer, larger tasks beat two workgroups per CU with smaller register slices).
constexpr int kDfMaxThreads = 320;

template <typename S, typename Acc, int R, int DPF>
__global__ void __launch_bounds__(kDfMaxThreads, 2) k_df_apply(DfArgs a) {
    constexpr int V = Raw<S>::V;
    typedef typename Raw<S>::T RawT;
    extern __shared__ unsigned char smem_raw[];
    Acc *xs = reinterpret_cast<Acc *>(smem_raw);          // maxdep * w
    Acc *red = xs + (size_t)a.maxdep * a.w;                // V*T
    __shared__ int sh_task, sh_abort;
    const int T = blockDim.x, tid = threadIdx.x, n = a.n;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int step = V * T;                                 // = dc * n
    int cl0[V];
#pragma unroll
    for (int v = 0; v < V; ++v) cl0[v] = (V * tid + v) / n;

    // each call moves the ticket counter by the task count plus one failing
    // grab per workgroup
    const unsigned long long base = a.calls * ((unsigned long long)a.ntasks + gridDim.x);
    auto grab = [&]() -> int {      // next ticket (lane 0)
        const unsigned long long t =
            __hip_atomic_fetch_add(a.ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - base;
        return (int)(t < (unsigned long long)a.ntasks ? t : (unsigned long long)a.ntasks);
    };
    if (wave == 0) {
        if (lane == 0) {
            sh_task = grab();
            sh_abort = 0;
        }
    }
    __syncthreads();
    for (;;) {
        const int t = __builtin_amdgcn_readfirstlane(sh_task);
        if (t >= a.ntasks) break;
        // the next ticket is taken now; its latency overlaps this task
        int next = 0;
        if (wave == 0 && lane == 0) next = grab();
        const int code = a.tasks[t];
        const int st = code & 3;
        const int rs = code >> 2;
        const int row = rs / a.S;
        const int sl = rs - row * a.S;
        const int c0 = sl * a.w;
        const int wc = min(n - c0, a.w);
        const int lim = wc * n;                             // flat elements of the slice
        int d0 = 0, nd = 1;
        const S *B;
        int64_t ld;
        if (st == ST_F) {
            d0 = a.lo_ptr[row]; nd = a.lo_ptr[row + 1] - d0;
            B = static_cast<const S *>(a.L); ld = a.ldl;
        } else if (st == ST_BA) {
            d0 = a.up_ptr[row]; nd = a.up_ptr[row + 1] - d0;
            B = static_cast<const S *>(a.U); ld = a.ldu;
        } else {
            B = static_cast<const S *>(a.Dinv); ld = a.ldl;
        }
        auto slot_of = [&](int d) -> int64_t {
            if (st == ST_F) return d0 + d;
            if (st == ST_BA) return a.up_idx[d0 + d];
            return row;
        };
        // 1. block slices into registers (they do not depend on other tasks)
        RawT raw[DPF][R];
#pragma unroll
        for (int d = 0; d < DPF; ++d) {
            const S *bp = d < nd ? B + slot_of(d) * ld + (size_t)c0 * n : B;
#pragma unroll
            for (int m = 0; m < R; ++m) {
                const int e = V * tid + m * step;
                if (d < nd && m < a.rounds && e < lim)
                    raw[d][m] = __builtin_nontemporal_load(reinterpret_cast<const RawT *>(bp + e));
                else
                    raw[d][m] = RawT{};
            }
        }
        // 2. wait until every row read by this task is complete
        if (wave == 0) {
            const int32_t *dep = nullptr;
            if (st == ST_F) {
                if (lane < nd) dep = a.cnt + a.lo_col[d0 + lane];
            } else if (st == ST_BA) {
                if (lane < nd) dep = a.cnt + 2 * a.nrows + a.up_col[d0 + lane];
            } else {
                if (lane == 0) dep = a.cnt + row;
                if (lane == 1) dep = a.cnt + a.nrows + row;
            }
            int spins = 0, abort = 0;
            for (;;) {
                const bool ok = dep == nullptr || ld_co(dep) >= a.target;
                if (__all(ok)) break;
                if (++spins > a.spin_limit) {
                    abort = 1;
                    if (lane == 0) __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                if ((spins & 255) == 0 && ld_co(a.err) != 0) { abort = 1; break; }
                __builtin_amdgcn_s_sleep(1);
            }
            // acquire: the partial loads below are ordered after the counter
            // observation (invalidates this CU's vector L1, not the L2)
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (lane == 0) sh_abort = abort;
        }
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(sh_abort)) break;
        // 3. the parents' slices into LDS
        for (int e = tid; e < nd * wc; e += T) {
            const int k = e / wc, c = e - k * wc;
            int parent = row;
            if (st == ST_F) parent = a.lo_col[d0 + k];
            else if (st == ST_BA) parent = a.up_col[d0 + k];
            xs[k * a.w + c] = static_cast<Acc>(df_view(a, st, parent, c0 + c));
        }
        __syncthreads();
        // 4. products
        Acc acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = 0;
#pragma unroll
        for (int d = 0; d < DPF; ++d) {
            if (d < nd) {
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    Acc x[V];
                    Raw<S>::get(raw[d][m], x);
                    const int e = V * tid + m * step;
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        if (e + v < lim) acc[v] += x[v] * xs[d * a.w + cl0[v] + m * a.dc];
                }
            }
        }
        for (int d = DPF; d < nd; ++d) {      // rows with more than DPF entries
            const S *bp = B + slot_of(d) * ld + (size_t)c0 * n;
            for (int m = 0; m < a.rounds; ++m) {
                const int e = V * tid + m * step;
                if (e < lim) {
                    Acc x[V];
                    Raw<S>::get(__builtin_nontemporal_load(reinterpret_cast<const RawT *>(bp + e)), x);
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        if (e + v < lim) acc[v] += x[v] * xs[d * a.w + cl0[v] + m * a.dc];
                }
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) red[V * tid + v] = acc[v];
        __syncthreads();
        // 5. partial out, then publish
        double *out = a.part + (size_t)st * a.plen + ((size_t)row * a.S + sl) * n;
        for (int r = tid; r < n; r += T) {
            Acc s = 0;
            for (int q = 0; q < a.dc; ++q) s += red[r + n * q];
            st_co(out + r, static_cast<double>(s));
        }
        // The partial was written with agent-scope atomic stores, which write
        // through this XCD's L2 to the device-coherent level: waiting for
        // their completion orders them before the counter increment.  (An
        // agent-scope release fence would also write back the whole L2 of
        // this XCD on every task, which serializes the sweep.)
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (wave == 0) {
            if (lane == 0) {
                __hip_atomic_fetch_add(a.cnt + (size_t)st * a.nrows + row, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                sh_task = next;
            }
        }
        __syncthreads();
    }
}

// z = sum_s PB[s]; all NaN when the sweep raised its error word
__global__ void k_df_finish(const double *__restrict__ pb, int S, int n, int64_t len, const int32_t *err,
                            double *__restrict__ z) {
    const bool bad = *err != 0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < len; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = e / n, c = e - row * n;
        double t = 0.0;
        for (int s = 0; s < S; ++s) t += pb[(row * S + s) * n + c];
        z[e] = bad ? __builtin_nan("") : t;
    }
}

__global__ void k_df_reset(int32_t *cnt, int64_t ncnt, unsigned long long *ticket, int32_t *err) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ncnt; e += (int64_t)gridDim.x * blockDim.x)
        cnt[e] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *err = 0;
}

// Registers per task: DPF block slices of at most R rounds (16 bytes per
// thread per round).
constexpr int kDPF = 3;
template <typename S> struct DfRounds;
template <> struct DfRounds<double> { static constexpr int R = 8; };
template <> struct DfRounds<float> { static constexpr int R = 6; };
template <> struct DfRounds<cdg_bf16> { static constexpr int R = 5; };

template <typename S, typename Acc>
const void *df_kernel() {
    return reinterpret_cast<const void *>(&k_df_apply<S, Acc, DfRounds<S>::R, kDPF>);
}

}  // namespace

// Geometry of the dataflow sweep for this block size and precision; fills
// M->df_* (not the device arrays).
int df_configure(cdg_ctx *ctx, cdg_prec *M) {
    const int prec = M->precision;
    const int V = prec == CDG_FP64 ? 2 : (prec == CDG_FP32 ? 4 : 8);
    const int Rmax = prec == CDG_FP64 ? DfRounds<double>::R : (prec == CDG_FP32 ? DfRounds<float>::R : DfRounds<cdg_bf16>::R);
    const StreamShape sh = stream_shape(M->n, V, 256, kDfMaxThreads);
    if (sh.T == 0) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "block size %d not supported by the sweep", M->n);
    M->df_T = sh.T;
    M->df_dc = sh.dc;
    const int cols_max = Rmax * sh.dc;
    int S = (M->n + cols_max - 1) / cols_max;
    const int per = (M->n + S - 1) / S;                     // balanced slice width
    M->df_rounds = (per + sh.dc - 1) / sh.dc;
    M->df_w = M->df_rounds * sh.dc;
    M->df_S = (M->n + M->df_w - 1) / M->df_w;
    const void *fn = prec == CDG_FP64 ? df_kernel<double, double>()
                   : (prec == CDG_FP32 ? df_kernel<float, float>() : df_kernel<cdg_bf16, float>());
    const size_t accsz = prec == CDG_FP64 ? 8 : 4;
    M->df_lds = accsz * ((size_t)M->maxdep * M->df_w + (size_t)V * sh.T);
    if (M->df_lds > 64 * 1024) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "sweep needs %zu bytes of LDS", M->df_lds);
    int per_cu = 0;
    CDG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, sh.T, M->df_lds));
    if (per_cu < 1) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "sweep kernel does not fit a CU");
    M->df_grid = per_cu * ctx->num_cus;
    // polls before a wait gives up (each poll sleeps ~64 cycles); CDG_DF_SPIN
    // lowers it, which is how the error path is exercised
    M->df_spin = 1 << 21;
    if (const char *e = std::getenv("CDG_DF_SPIN")) M->df_spin = std::max(1, std::atoi(e));
    return CDG_OK;
}


int df_reset(cdg_ctx *ctx, cdg_prec *M) {
    const int64_t ncnt = 3 * (int64_t)M->nrows;
    hipLaunchKernelGGL(k_df_reset, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((ncnt + 255) / 256, 1024))),
                       dim3(256), 0, ctx->stream, M->d_cnt, ncnt, M->d_ticket, M->d_dferr);
    CDG_HIP_TRY(hipGetLastError());
    M->df_epoch = 0;
    return CDG_OK;
}

namespace {

template <typename S, typename Acc>
int df_launch(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    if (M->df_epoch >= (1 << 30) / std::max(1, M->df_S)) {   // keep epoch*S inside int32
        int e = df_reset(ctx, M);
        if (e) return e;
    }
    DfArgs a{};
    a.L = M->L;
    a.Dinv = M->Dinv;
    a.ldl = M->ld;
    if (M->precision == CDG_FP64) { a.U = A->blocks; a.ldu = A->ld; }
    else { a.U = M->U; a.ldu = M->ld; }
    a.lo_ptr = M->d_lo_ptr; a.lo_col = M->d_lo_col;
    a.up_ptr = M->d_up_ptr; a.up_idx = M->d_up_idx; a.up_col = M->d_up_col;
    a.tasks = M->d_tasks;
    a.ntasks = M->df_ntasks;
    a.cnt = M->d_cnt;
    a.ticket = M->d_ticket;
    const int grid = std::min(M->df_grid, M->df_ntasks);
    a.calls = (unsigned long long)M->df_epoch;
    a.err = M->d_dferr;
    a.v = v;
    a.part = M->d_part;
    a.plen = (int64_t)M->nrows * M->df_S * M->n;
    a.nrows = M->nrows; a.n = M->n; a.S = M->df_S; a.w = M->df_w; a.rounds = M->df_rounds; a.dc = M->df_dc;
    a.maxdep = M->maxdep;
    a.spin_limit = M->df_spin;
    a.target = (M->df_epoch + 1) * M->df_S;
    hipLaunchKernelGGL((k_df_apply<S, Acc, DfRounds<S>::R, kDPF>), dim3(grid), dim3(M->df_T), M->df_lds, ctx->stream, a);
    CDG_HIP_TRY(hipGetLastError());
    M->df_epoch += 1;

    const int64_t len = (int64_t)M->nrows * M->n;
    hipLaunchKernelGGL(k_df_finish, dim3((unsigned)std::min<int64_t>((len + 255) / 256, 4096)), dim3(256), 0,
                       ctx->stream, M->d_part + 2 * a.plen, M->df_S, M->n, len, M->d_dferr, z);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

}  // namespace

int launch_ilu_apply_df(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    if (M->nrows == 0) return CDG_OK;
    if (M->precision == CDG_FP64) return df_launch<double, double>(ctx, M, A, v, z);
    if (M->precision == CDG_FP32) return df_launch<float, float>(ctx, M, A, v, z);
    return df_launch<cdg_bf16, float>(ctx, M, A, v, z);
}
