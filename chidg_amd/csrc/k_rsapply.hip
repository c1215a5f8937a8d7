// ILU(0) apply as one persistent sweep over row slices (precon_ILU0.f90:286-411).
//
//   forward   y_i = v_i - sum_{j<i} L_ij y_j                 (precon_ILU0.f90:336-359)
//   backward  z_i = Dinv_i y_i - sum_{k>i} W_ik z_k,  W_ik = Dinv_i U_ik
//             (the reference's z_i -= U_ik z_k followed by z_i = Dinv_i z_i,
//              precon_ILU0.f90:363-398, with Dinv_i multiplied in beforehand)
//
// Storage ("row-slice" layout, written once per factor by the fold and
// relayout kernels): every n x n block is cut into S = n / w slices of w
// rows; slice s holds rows s w .. s w + w - 1 of all n columns, column-major
// with leading dimension w, at element offset s * SL of the block's slot
// (SL = w n rounded up to 8 elements, so every slice starts 16-byte aligned).
// A task (row i, slice s) therefore streams one contiguous range per block
// and produces final values of rows s w .. s w + w - 1 of y_i or z_i: no
// partial sums travel between tasks.
//
// Scheduling: tasks are listed in a topological order (forward levels, then
// backward levels).  Task t belongs to class t mod NC (NC <= 8, one class per
// XCD when the grid is dealt round-robin over the XCDs); workgroup g serves
// class g mod NC and takes that class's tasks in list order from the class's
// ticket counter (one agent-scope atomic per task, fetched two tasks ahead).
// Tickets go out in order to running workgroups only, so the sweep makes
// progress whatever the number of resident workgroups: the lowest unfinished
// task has finished parents and is held, or is next to be taken, by a running
// workgroup of its class.  (A fixed round-robin deal instead needs every
// workgroup resident at once, which the occupancy figure does not guarantee.)
// Two counter sets alternate between calls: workgroup 0 clears the set the
// next call uses.  Each wait is bounded nevertheless: past the limit a
// workgroup raises the error word; every workgroup that sees it drains its
// class's remaining tickets without waiting, writing NaN into the outputs of
// backward tasks, so the sweep always ends.
//
// Exchange between tasks: every value of y and z sits in a 64-bit word
// together with a 32-bit tag, the number of the apply call that wrote it
// (FP64 values take two words, each half with the tag).  A consumer loads the
// parent words with device-coherent (agent-scope) atomic loads and accepts
// them once every tag is the current call's; a 64-bit aligned access is
// single-copy atomic, so value and tag are seen together.  No completion
// counters, no fences, nothing to reset between calls but the ticket set.
//
// Software pipeline per workgroup: the block slices of the next task are
// loaded into registers while the current task reduces and stores, so a task
// that waits for its parents already holds its blocks.
#include "cdg_internal.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

namespace {

typedef double rs_f64x2 __attribute__((ext_vector_type(2)));
typedef float rs_f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t rs_u32x4 __attribute__((ext_vector_type(4)));

// 16-byte raw vector of each storage type, its V elements and their
// accumulation type
template <typename S> struct Raw16;
template <> struct Raw16<double> {
    typedef rs_f64x2 T;
    typedef double Acc;
    static constexpr int V = 2;
    __device__ static __forceinline__ void get(const T &r, double (&o)[2]) { o[0] = r.x; o[1] = r.y; }
};
template <> struct Raw16<float> {
    typedef rs_f32x4 T;
    typedef float Acc;
    static constexpr int V = 4;
    __device__ static __forceinline__ void get(const T &r, float (&o)[4]) {
        o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = r.w;
    }
};
template <> struct Raw16<cdg_bf16> {
    typedef rs_u32x4 T;
    typedef float Acc;
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

__device__ __forceinline__ uint64_t ld_tag(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_tag(uint64_t *p, uint64_t w) {
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tagged value formats: Words 64-bit words per value.
template <typename VT> struct Tagged;
template <> struct Tagged<float> {
    static constexpr int Words = 1;
    __device__ static __forceinline__ void put(uint64_t *p, double v, uint32_t tag) {
        st_tag(p, ((uint64_t)tag << 32) | __float_as_uint(static_cast<float>(v)));
    }
    // value of the loaded words; ok = every word carries `tag`
    __device__ static __forceinline__ float get(const uint64_t (&w)[1], uint32_t tag, bool &ok) {
        ok = (uint32_t)(w[0] >> 32) == tag;
        return __uint_as_float((uint32_t)w[0]);
    }
};
template <> struct Tagged<double> {
    static constexpr int Words = 2;
    __device__ static __forceinline__ void put(uint64_t *p, double v, uint32_t tag) {
        const uint64_t b = (uint64_t)__double_as_longlong(v);
        st_tag(p, ((uint64_t)tag << 32) | (b & 0xffffffffull));
        st_tag(p + 1, ((uint64_t)tag << 32) | (b >> 32));
    }
    __device__ static __forceinline__ double get(const uint64_t (&w)[2], uint32_t tag, bool &ok) {
        ok = (uint32_t)(w[0] >> 32) == tag && (uint32_t)(w[1] >> 32) == tag;
        return __longlong_as_double((long long)((w[0] & 0xffffffffull) | (w[1] << 32)));
    }
};

// Row sums of a task: red[r + w q] (q < dc) holds the dc partial products of
// row r; K = min(T / w, dc) groups of w threads first sum every K-th partial,
// then thread r < w sums the K group sums (fixed order: the result does not
// depend on timing).  Returns row tid's sum for tid < w.  Ends with the
// partials consumed (one barrier inside).
template <typename Acc>
__device__ __forceinline__ Acc rs_rowsum(const Acc *red, Acc *red2, int w, int dc, int T, int tid) {
    const int K = max(1, min(T / w, dc));
    if (tid < K * w) {
        const int g = tid / w, r = tid - g * w;
        Acc s = 0;
        for (int q = g; q < dc; q += K) s += red[r + w * q];
        red2[g * w + r] = s;
    }
    __syncthreads();
    Acc s = 0;
    if (tid < w)
        for (int g = 0; g < K; ++g) s += red2[g * w + tid];
    return s;
}

struct RsArgs {
    const int32_t *tasks;     // kRsRec ints per task (see rs_build_tasks)
    int32_t ntasks;
    const void *Sb;           // row-slice slots: nlo L, nup W, nrows Dinv (one allocation)
    int64_t slot;             // elements per slot
    int32_t SL;               // elements per slice
    int32_t n, w, dc;         // block size, rows per slice, columns per walk step
    const double *v;          // the vector preconditioned
    double *z;                // the result
    uint64_t *yb, *zb;        // tagged forward / backward results
    uint32_t tag;
    uint32_t *ctl;            // control words: kRsCtlWords (error word, ticket counters)
    int32_t nclass;           // ticket classes NC (<= 8)
    int32_t spin_limit;
};

}  // namespace

// Task record: [0] row, [1] slice | nblocks << 8 | backward << 16,
// [2 + k] slot of block k in the row-slice storage (L slots 0..nlo-1, W
// slots nlo..nlo+nup-1, Dinv slots after them), [2 + kRsMaxBlocks + k] parent
// row of block k.  Backward tasks: block 0 is the row's Dinv (parent: the
// row's own y), blocks 1.. the W of its upper entries (parents: their z).
constexpr int kRsMaxBlocks = 7;
constexpr int kRsRec = 16;
constexpr int kRsMaxThreads = 512;

namespace {

// Control words of a sweep (uint32): [0] error word; the ticket counter of
// set s (0, 1) and class c < kRsMaxClass at kRsLine * (1 + kRsMaxClass s + c),
// one 128-byte line each.
constexpr int kRsLine = 32;
constexpr int kRsMaxClass = 256;
constexpr int kRsCtlWords = kRsLine * (1 + 2 * kRsMaxClass);

__device__ __forceinline__ uint32_t *rs_counter(uint32_t *ctl, uint32_t set, int cls) {
    return ctl + kRsLine * (1 + kRsMaxClass * (set & 1u) + cls);
}

template <typename S, int NB, int R, int EPT>
__global__ void __launch_bounds__(kRsMaxThreads) k_rs_sweep(RsArgs args) {
    typedef Raw16<S> RW;
    typedef typename RW::Acc Acc;
    typedef typename RW::T RawT;
    typedef typename std::conditional<std::is_same<S, double>::value, double, float>::type VT;
    typedef Tagged<VT> TG;
    constexpr int V = RW::V, WD = TG::Words;
    // the arguments as scalars (no lambda refers to the argument block, which
    // would otherwise be copied to per-lane scratch memory)
    const int32_t *const tasks = args.tasks;
    const int ntasks = args.ntasks;
    const S *const Sb = static_cast<const S *>(args.Sb);
    const int64_t slot = args.slot;
    const int SL = args.SL, n = args.n, w = args.w, dc = args.dc;
    const double *const vin = args.v;
    double *const zout = args.z;
    uint64_t *const yb = args.yb;
    uint64_t *const zb = args.zb;
    const uint32_t tag = args.tag;
    uint32_t *const ctl = args.ctl;
    const int NC = args.nclass, spin_limit = args.spin_limit;

    extern __shared__ unsigned char smem_raw[];
    __shared__ int s_tick[2];
    const int T = blockDim.x, tid = threadIdx.x;
    Acc *xs = reinterpret_cast<Acc *>(smem_raw);                    // kRsMaxBlocks * n
    Acc *red = xs + (((size_t)kRsMaxBlocks * n + 3) & ~(size_t)3);  // V*T
    Acc *red2 = red + (size_t)V * blockDim.x;                       // T
    const int lim = w * n;             // elements of one block slice
    const int step = V * T;            // = dc * w
    const int lim_ld = (lim - 1) / V * V;   // last 16-byte group starting inside the slice
    int col0[V];
#pragma unroll
    for (int v = 0; v < V; ++v) col0[v] = (V * tid + v) / w;
    // w >= V: a 16-byte group spans at most two columns, `split` values in the first
    const bool two = w >= V;
    const int split = min(V, w * (col0[0] + 1) - V * tid);

    const int cls = blockIdx.x % NC;
    uint32_t *const cnt = rs_counter(ctl, tag, cls);
    // workgroup 0 clears the counters of the next call, all of them (the next
    // call may have more classes); the previous call, which used them, has
    // ended: calls on one buffer set are stream-ordered
    if (blockIdx.x == 0)
        for (int c = tid; c < kRsMaxClass; c += T)
            __hip_atomic_store(rs_counter(ctl, tag + 1u, c), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // ticket k of the class is task k NC + cls
    auto ticket = [cnt, NC, cls]() -> int {
        const uint32_t k = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return (int)min(k, (uint32_t)INT32_MAX / 16u) * NC + cls;
    };
    if (tid == 0) {
        s_tick[0] = ticket();
        s_tick[1] = ticket();
    }
    __syncthreads();
    int t = s_tick[0], tn = s_tick[1];
    __syncthreads();
    if (t >= ntasks) return;

    auto slice_base = [Sb, slot, SL](const int32_t *rec, int k) -> const S * {
        const int nb = (rec[1] >> 8) & 0xff, sl = rec[1] & 0xff;
        const int kk = k < nb ? k : 0;
        return Sb + (size_t)rec[2 + kk] * slot + (size_t)sl * SL;
    };
    RawT raw[NB][R];
    // Every load is issued, none under a branch (a branch here would make the
    // compiler wait for all outstanding loads at the join): a block the task
    // does not have reads the first 16 bytes of block 0, the same address in
    // every lane (one request per wave), and steps past the slice re-read its
    // last 16-byte group.
    auto issue = [&raw, &slice_base, tid, step, lim_ld](const int32_t *rec) {
        const int nb = (rec[1] >> 8) & 0xff;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const S *b = slice_base(rec, k);
#pragma unroll
            for (int m = 0; m < R; ++m)
                raw[k][m] = __builtin_nontemporal_load(
                    reinterpret_cast<const RawT *>(b + (k < nb ? min(V * tid + m * step, lim_ld) : 0)));
        }
    };
    issue(tasks + (size_t)t * kRsRec);
    int aborted = 0;
    int tnn = 0;
    for (;;) {
        const int32_t *rec = tasks + (size_t)t * kRsRec;
        const int row = rec[0];
        const int sl = rec[1] & 0xff, nb = (rec[1] >> 8) & 0xff, bwd = (rec[1] >> 16) & 1;
        // the ticket after next goes out now; it is read after the row sums
        if (tid == 0) tnn = ticket();
        // ---- 1. parents into LDS: nb vectors of n values, all tags current
        // (skipped once the sweep is abandoned: the workgroup then only
        // drains its class's tickets, writing NaN for backward tasks)
        if (!aborted) {
            // element j = k n + c (thread's j = tid + i T) is value c of parent
            // k: word offset off[i] in y (forward parents and the backward
            // row's own y, j < n) or z (backward upper parents, j >= n); the
            // backward row's own y enters with +, all others with -
            const int ne = nb * n;
            // the parents are wave-uniform (scalar loads); each lane picks its
            // own by selects, not by a per-lane load from the record
            int par[kRsMaxBlocks];
#pragma unroll
            for (int k = 0; k < kRsMaxBlocks; ++k) par[k] = rec[2 + kRsMaxBlocks + k];
            int off[EPT];
#pragma unroll
            for (int i = 0; i < EPT; ++i) {
                // clamped into [0, ne): a task without parents (ne = 0) reads
                // value 0 of row 0, unused
                const int j = max(0, min(tid + i * T, ne - 1));
                const int k = j / n, c = j - k * n;
                int pk = par[0];
#pragma unroll
                for (int q = 1; q < kRsMaxBlocks; ++q) pk = k == q ? par[q] : pk;
                off[i] = (pk * n + c) * WD;
            }
            uint32_t need = 0;
#pragma unroll
            for (int i = 0; i < EPT; ++i)
                if (tid + i * T < ne) need |= 1u << i;
            int spins = 0;
            for (;;) {
                uint64_t wd[EPT][WD];
#pragma unroll
                for (int i = 0; i < EPT; ++i) {
                    const uint64_t *src = (bwd && tid + i * T >= n ? zb : yb) + off[i];
#pragma unroll
                    for (int q = 0; q < WD; ++q) wd[i][q] = ld_tag(src + q);
                }
#pragma unroll
                for (int i = 0; i < EPT; ++i) {
                    bool ok;
                    const VT x = TG::get(wd[i], tag, ok);
                    if (ok && ((need >> i) & 1u)) {
                        const int j = tid + i * T;
                        xs[j] = (bwd && j < n) ? static_cast<Acc>(x) : -static_cast<Acc>(x);
                        need &= ~(1u << i);
                    }
                }
                if (!__syncthreads_or(need != 0)) break;
                // bounded wait; the error word, set to this call's tag by the
                // first workgroup that gives up, ends every other one's wait (a
                // word left by an earlier call holds another tag and stops nothing)
                ++spins;
                int stop = spins > spin_limit;
                if (stop && tid == 0) __hip_atomic_store(ctl, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((spins & 63) == 0) stop |= __hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == tag;
                if (__syncthreads_or(stop)) { aborted = 1; break; }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        // ---- 2. products: this thread's elements of the nb block slices
        Acc acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = 0;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (k < nb && !aborted) {
                const Acc *xk = xs + k * n;
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    const int e = V * tid + m * step;
                    Acc x[V];
                    RW::get(raw[k][m], x);
                    if (two) {   // the group's values in two columns: two LDS reads
                        const Acc xa = xk[min(col0[0] + m * dc, n - 1)], xb = xk[min(col0[0] + 1 + m * dc, n - 1)];
#pragma unroll
                        for (int v = 0; v < V; ++v)
                            if (e + v < lim) acc[v] += x[v] * (v < split ? xa : xb);
                    } else {
#pragma unroll
                        for (int v = 0; v < V; ++v)
                            if (e + v < lim) acc[v] += x[v] * xk[min(col0[v] + m * dc, n - 1)];
                    }
                }
            }
        }
        // slices longer than R steps, rows with more than NB blocks: the rest from memory
        if ((R * step < lim || nb > NB) && !aborted) {
            for (int k = 0; k < nb; ++k) {
                const S *b = slice_base(rec, k);
                const Acc *xk = xs + k * n;
                for (int m = k < NB ? R : 0; m * step < lim; ++m) {
                    const int e = V * tid + m * step;
                    if (e < lim) {
                        Acc x[V];
                        RW::get(__builtin_nontemporal_load(reinterpret_cast<const RawT *>(b + e)), x);
#pragma unroll
                        for (int v = 0; v < V; ++v)
                            if (e + v < lim) acc[v] += x[v] * xk[min(col0[v] + m * dc, n - 1)];
                    }
                }
            }
        }
        // ---- 3. the next task's blocks go out now (registers are free)
        if (tn < ntasks && !aborted) issue(tasks + (size_t)tn * kRsRec);
        // ---- 4. row sums (fixed order) and the outputs of this task
#pragma unroll
        for (int v = 0; v < V; ++v) red[V * tid + v] = acc[v];
        if (tid == 0) s_tick[0] = tnn;
        __syncthreads();
        const Acc s = rs_rowsum(red, red2, w, dc, T, tid);
        if (tid < w) {
            const size_t gi = (size_t)row * n + (size_t)sl * w + tid;
            if (aborted) {
                if (bwd) zout[gi] = __builtin_nan("");
            } else if (!bwd) {
                const double y = vin[gi] + static_cast<double>(s);
                TG::put(yb + gi * WD, y, tag);
            } else {
                const VT zz = static_cast<VT>(s);
                TG::put(zb + gi * WD, static_cast<double>(zz), tag);
                zout[gi] = static_cast<double>(zz);
            }
        }
        tnn = s_tick[0];
        __syncthreads();
        t = tn;
        tn = tnn;
        if (t >= ntasks) break;
    }
}

// ---- the same tasks, one launch per wavefront level ------------------------------
// Level l of the forward sweep is tasks [lo, hi) of the task list, and so is
// each backward level: every task of a launch has its parents in earlier
// launches (stream order makes them visible), so a task loads its parents
// with plain loads, streams its block slices and writes final values of its
// rows -- no tags, no waits, no partial sums.  y (the forward result) is kept
// in the storage's accumulation precision, as the persistent sweep keeps it.
struct RslArgs {
    const int32_t *tasks;
    int32_t t0;               // first task of the level
    const void *Sb;
    int64_t slot;
    int32_t SL, n, w, dc;
    const double *v;
    double *z;
    void *y;                  // nrows * n values of the accumulation type
};

template <typename S, int NB, int R, int EPT>
__global__ void __launch_bounds__(kRsMaxThreads) k_rs_level(RslArgs args) {
    typedef Raw16<S> RW;
    typedef typename RW::Acc Acc;
    typedef typename RW::T RawT;
    typedef typename std::conditional<std::is_same<S, double>::value, double, float>::type VT;
    constexpr int V = RW::V;
    const S *const Sb = static_cast<const S *>(args.Sb);
    const int64_t slot = args.slot;
    const int SL = args.SL, n = args.n, w = args.w, dc = args.dc;
    VT *const y = static_cast<VT *>(args.y);
    double *const zout = args.z;
    extern __shared__ unsigned char smem_raw[];
    const int T = blockDim.x, tid = threadIdx.x;
    Acc *xs = reinterpret_cast<Acc *>(smem_raw);
    Acc *red = xs + (((size_t)kRsMaxBlocks * n + 3) & ~(size_t)3);
    Acc *red2 = red + (size_t)V * blockDim.x;
    const int lim = w * n, step = V * T, lim_ld = (lim - 1) / V * V;
    const int32_t *rec = args.tasks + (size_t)(args.t0 + (int)blockIdx.x) * kRsRec;
    const int row = rec[0];
    const int sl = rec[1] & 0xff, nb = (rec[1] >> 8) & 0xff, bwd = (rec[1] >> 16) & 1;
    // parents first (one batch of loads, clamped in bounds, no branch), then
    // the block slices; the parents' wait then leaves the block loads in flight
    const int ne = nb * n;
    int par[kRsMaxBlocks];
#pragma unroll
    for (int k = 0; k < kRsMaxBlocks; ++k) par[k] = rec[2 + kRsMaxBlocks + k];
    VT pv[EPT];
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
        const int j = max(0, min(tid + i * T, ne - 1));
        const int k = j / n, c = j - k * n;
        int pk = par[0];
#pragma unroll
        for (int q = 1; q < kRsMaxBlocks; ++q) pk = k == q ? par[q] : pk;
        const size_t gi = (size_t)pk * n + c;
        // forward parents and the backward row's own value: y; backward
        // upper parents: z (stored as double, holding VT values)
        pv[i] = (bwd && j >= n) ? static_cast<VT>(zout[gi]) : y[gi];
    }
    RawT raw[NB][R];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int kk = k < nb ? k : 0;
        const S *b = Sb + (size_t)rec[2 + kk] * slot + (size_t)sl * SL;
#pragma unroll
        for (int m = 0; m < R; ++m)
            raw[k][m] = __builtin_nontemporal_load(
                reinterpret_cast<const RawT *>(b + (k < nb ? min(V * tid + m * step, lim_ld) : 0)));
    }
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
        const int j = tid + i * T;
        if (j < ne) xs[j] = (bwd && j < n) ? static_cast<Acc>(pv[i]) : -static_cast<Acc>(pv[i]);
    }
    __syncthreads();
    int col0[V];
#pragma unroll
    for (int v = 0; v < V; ++v) col0[v] = (V * tid + v) / w;
    // w >= V: a 16-byte group spans at most two columns, `split` values in the first
    const bool two = w >= V;
    const int split = min(V, w * (col0[0] + 1) - V * tid);
    Acc acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (k < nb) {
            const Acc *xk = xs + k * n;
#pragma unroll
            for (int m = 0; m < R; ++m) {
                const int e = V * tid + m * step;
                Acc x[V];
                RW::get(raw[k][m], x);
                if (two) {   // the group's values in two columns: two LDS reads
                    const Acc xa = xk[min(col0[0] + m * dc, n - 1)], xb = xk[min(col0[0] + 1 + m * dc, n - 1)];
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        if (e + v < lim) acc[v] += x[v] * (v < split ? xa : xb);
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        if (e + v < lim) acc[v] += x[v] * xk[min(col0[v] + m * dc, n - 1)];
                }
            }
        }
    }
    if (R * step < lim || nb > NB) {
        for (int k = 0; k < nb; ++k) {
            const S *b = Sb + (size_t)rec[2 + k] * slot + (size_t)sl * SL;
            const Acc *xk = xs + k * n;
            for (int m = k < NB ? R : 0; m * step < lim; ++m) {
                const int e = V * tid + m * step;
                if (e < lim) {
                    Acc x[V];
                    RW::get(__builtin_nontemporal_load(reinterpret_cast<const RawT *>(b + e)), x);
                    if (two) {   // the group's values in two columns: two LDS reads
                        const Acc xa = xk[min(col0[0] + m * dc, n - 1)], xb = xk[min(col0[0] + 1 + m * dc, n - 1)];
#pragma unroll
                        for (int v = 0; v < V; ++v)
                            if (e + v < lim) acc[v] += x[v] * (v < split ? xa : xb);
                    } else {
#pragma unroll
                        for (int v = 0; v < V; ++v)
                            if (e + v < lim) acc[v] += x[v] * xk[min(col0[v] + m * dc, n - 1)];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) red[V * tid + v] = acc[v];
    __syncthreads();
    const Acc s = rs_rowsum(red, red2, w, dc, T, tid);
    if (tid < w) {
        const size_t gi = (size_t)row * n + (size_t)sl * w + tid;
        if (!bwd)
            y[gi] = static_cast<VT>(args.v[gi] + static_cast<double>(s));
        else
            zout[gi] = static_cast<double>(static_cast<VT>(s));
    }
}

// ---- k right-hand sides per task ------------------------------------------------
// The row-slice level launches of k vectors at once (v, z and y hold k
// vectors `stride` elements apart): a task streams its block slices once and
// applies them to all k vectors.  Each vector's products, sums and roundings
// are the single-vector kernel's, so each gets the same bits as alone.
struct RsmArgs {
    const int32_t *tasks;
    int32_t t0;
    const void *Sb;
    int64_t slot;
    int32_t SL, n, w, dc;
    const double *v;
    double *z;
    void *y;                  // k vectors of nrows * n values of the accumulation type
    int64_t stride;
    int32_t k;                // right-hand sides of this launch (<= K)
    int32_t nbmax;            // blocks per task at most (parents per task)
    const int32_t *row_k;     // null, or right-hand sides of each row's system (a batch of
                              // systems in one matrix): this launch serves r0 .. r0+K-1
    int32_t r0;
};

// FP64/FP32 with K = 8 and R <= 4 need a little over 128 VGPRs unconstrained,
// which leaves one 512-thread workgroup per CU; held to 128 two fit, and
// one's loads overlap the other's products and row sums (BF16 would spill).
template <typename S, int K, int R>
__global__ void __launch_bounds__(kRsMaxThreads)
__attribute__((amdgpu_waves_per_eu((K >= 8 && R <= 4 && sizeof(S) >= 4) ? 4 : 1)))
k_rs_multi(RsmArgs args) {
    typedef Raw16<S> RW;
    typedef typename RW::Acc Acc;
    typedef typename RW::T RawT;
    typedef typename std::conditional<std::is_same<S, double>::value, double, float>::type VT;
    constexpr int V = RW::V, NB = 4;
    const S *const Sb = static_cast<const S *>(args.Sb);
    const int64_t slot = args.slot, stride = args.stride;
    const int SL = args.SL, n = args.n, w = args.w, dc = args.dc, nbmax = args.nbmax;
    VT *const y = static_cast<VT *>(args.y);
    double *const zout = args.z;
    extern __shared__ unsigned char smem_raw[];
    const int T = blockDim.x, tid = threadIdx.x;
    constexpr int R2 = K < 4 ? K : (V >= 8 ? 2 : 4);                 // right-hand sides per row-sum round
    const int xsz = nbmax * n;                                       // parents of one right-hand side
    Acc *xs = reinterpret_cast<Acc *>(smem_raw);                     // K * xsz
    Acc *red = xs;                                                   // R2 * V*T, over xs once the products are done
    Acc *red2 = xs + max((((size_t)K * xsz + 3) & ~(size_t)3), (size_t)R2 * V * T);   // R2 * T
    const int lim = w * n, step = V * T, lim_ld = (lim - 1) / V * V;
    const int32_t *rec = args.tasks + (size_t)(args.t0 + (int)blockIdx.x) * kRsRec;
    const int row = rec[0];
    const int k = args.row_k ? min(max(args.row_k[row] - args.r0, 0), K) : args.k;
    if (k == 0) return;
    const int sl = rec[1] & 0xff, nb = (rec[1] >> 8) & 0xff, bwd = (rec[1] >> 16) & 1;
    // the block slices first (they do not depend on the parents), then the
    // parents of all k right-hand sides into LDS
    RawT raw[NB][R];
#pragma unroll
    for (int kb = 0; kb < NB; ++kb) {
        const int kk = kb < nb ? kb : 0;
        const S *b = Sb + (size_t)rec[2 + kk] * slot + (size_t)sl * SL;
#pragma unroll
        for (int m = 0; m < R; ++m)
            raw[kb][m] = __builtin_nontemporal_load(
                reinterpret_cast<const RawT *>(b + (kb < nb ? min(V * tid + m * step, lim_ld) : 0)));
    }
    // each parent value's offset once, then its k loads together (no load
    // waits on the previous one's LDS store)
    const int ne = nb * n;
    for (int jj = tid; jj < ne; jj += T) {
        const int kb = jj / n, c = jj - kb * n;
        const size_t gb = (size_t)rec[2 + kRsMaxBlocks + kb] * n + c;
        const bool fromz = bwd && jj >= n, own = bwd && jj < n;
        VT x[K];
#pragma unroll
        for (int r = 0; r < K; ++r)
            if (r < k) x[r] = fromz ? static_cast<VT>(zout[(size_t)r * stride + gb]) : y[(size_t)r * stride + gb];
#pragma unroll
        for (int r = 0; r < K; ++r)
            if (r < k) xs[r * xsz + jj] = own ? static_cast<Acc>(x[r]) : -static_cast<Acc>(x[r]);
    }
    __syncthreads();
    int col0[V];
#pragma unroll
    for (int v = 0; v < V; ++v) col0[v] = (V * tid + v) / w;
    const bool two = w >= V;
    const int split = min(V, w * (col0[0] + 1) - V * tid);
    Acc acc[K][V];
#pragma unroll
    for (int r = 0; r < K; ++r)
#pragma unroll
        for (int v = 0; v < V; ++v) acc[r][v] = 0;
    auto fma_group = [&](const Acc (&x)[V], int e, int kb, int m) {
#pragma unroll
        for (int r = 0; r < K; ++r) {
            const Acc *xk = xs + r * xsz + kb * n;
            if (two) {
                const Acc xa = xk[min(col0[0] + m * dc, n - 1)], xb = xk[min(col0[0] + 1 + m * dc, n - 1)];
#pragma unroll
                for (int v = 0; v < V; ++v)
                    if (e + v < lim) acc[r][v] += x[v] * (v < split ? xa : xb);
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v)
                    if (e + v < lim) acc[r][v] += x[v] * xk[min(col0[v] + m * dc, n - 1)];
            }
        }
    };
#pragma unroll
    for (int kb = 0; kb < NB; ++kb) {
        if (kb < nb) {
#pragma unroll
            for (int m = 0; m < R; ++m) {
                Acc x[V];
                RW::get(raw[kb][m], x);
                fma_group(x, V * tid + m * step, kb, m);
            }
        }
    }
    if (R * step < lim || nb > NB) {
        for (int kb = 0; kb < nb; ++kb) {
            const S *b = Sb + (size_t)rec[2 + kb] * slot + (size_t)sl * SL;
            for (int m = kb < NB ? R : 0; m * step < lim; ++m) {
                const int e = V * tid + m * step;
                if (e < lim) {
                    Acc x[V];
                    RW::get(__builtin_nontemporal_load(reinterpret_cast<const RawT *>(b + e)), x);
                    fma_group(x, e, kb, m);
                }
            }
        }
    }
    // row sums, R2 right-hand sides per round (each in rs_rowsum's order)
    const int G = max(1, min(T / w, dc));
#pragma unroll
    for (int r0 = 0; r0 < K; r0 += R2) {
        if (r0 >= k) break;
        __syncthreads();   // the products are done with xs; the previous round with red and red2
#pragma unroll
        for (int q = 0; q < R2; ++q)
            if (r0 + q < k)
#pragma unroll
                for (int v = 0; v < V; ++v) red[(size_t)q * step + V * tid + v] = acc[r0 + q][v];
        __syncthreads();
        if (tid < G * w) {
            const int g = tid / w, rr = tid - g * w;
#pragma unroll
            for (int q = 0; q < R2; ++q) {
                if (r0 + q < k) {
                    const Acc *rq = red + (size_t)q * step;
                    Acc s = 0;
                    for (int p = g; p < dc; p += G) s += rq[rr + w * p];
                    red2[q * T + g * w + rr] = s;
                }
            }
        }
        __syncthreads();
        if (tid < w) {
            const size_t gi = (size_t)row * n + (size_t)sl * w + tid;
#pragma unroll
            for (int q = 0; q < R2; ++q) {
                const int r = r0 + q;
                if (r < k) {
                    Acc s = 0;
                    for (int g = 0; g < G; ++g) s += red2[q * T + g * w + tid];
                    if (!bwd)
                        y[(size_t)r * stride + gi] =
                            static_cast<VT>(args.v[(size_t)r * stride + gi] + static_cast<double>(s));
                    else
                        zout[(size_t)r * stride + gi] = static_cast<double>(static_cast<VT>(s));
                }
            }
        }
    }
}

// ---- conversions to the storage type ----------------------------------------
template <typename SO> __device__ __forceinline__ SO to_store(double v);
template <> __device__ __forceinline__ double to_store<double>(double v) { return v; }
template <> __device__ __forceinline__ float to_store<float>(double v) { return static_cast<float>(v); }
template <> __device__ __forceinline__ cdg_bf16 to_store<cdg_bf16>(double v) {
    uint32_t u = __float_as_uint(static_cast<float>(v));
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return static_cast<cdg_bf16>(u >> 16);
}

// Column-major n x n slots -> row-slice slots (one workgroup per block).
template <typename SI, typename SO>
__global__ void __launch_bounds__(256) k_to_rs(const SI *__restrict__ src, int64_t ld_in, SO *__restrict__ dst,
                                               int64_t slot, int count, int n, int w, int SL) {
    const int nn = n * n;
    for (int b = blockIdx.x; b < count; b += gridDim.x) {
        const SI *s = src + (size_t)b * ld_in;
        SO *d = dst + (size_t)b * slot;
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const int c = e / n, r = e - c * n, q = r / w;
            d[(size_t)q * SL + (size_t)c * w + (r - q * w)] = to_store<SO>(static_cast<double>(s[e]));
        }
    }
}

}  // namespace

// ---- host side ------------------------------------------------------------------
// Rows per slice: a divisor of n (equal slices) no larger than wmax.
int rs_slice_rows(int n, int wmax) {
    int best = 1;
    for (int d = 1; d <= n; ++d)
        if (n % d == 0 && d <= wmax) best = d;
    return best;
}

namespace {

template <typename S>
int R_needed(int lim, int step) { return (lim + step - 1) / step; }

// kernel table: S x R in {2, 3, 4, 6, 8} x EPT in {2, 4, 8}, NB = 4
template <typename S>
const void *rs_kernel(int R, int EPT) {
#define CDG_RS_K(RR, EE) reinterpret_cast<const void *>(&k_rs_sweep<S, 4, RR, EE>)
#define CDG_RS_E(RR)                                  \
    if (EPT <= 2) return CDG_RS_K(RR, 2);             \
    if (EPT <= 4) return CDG_RS_K(RR, 4);             \
    return CDG_RS_K(RR, 8);
    if (R <= 2) { CDG_RS_E(2) }
    if (R <= 3) { CDG_RS_E(3) }
    if (R <= 4) { CDG_RS_E(4) }
    if (R <= 6) { CDG_RS_E(6) }
    CDG_RS_E(8)
#undef CDG_RS_E
#undef CDG_RS_K
}

template <typename S>
const void *rsl_kernel(int R, int EPT) {
#define CDG_RS_K(RR, EE) reinterpret_cast<const void *>(&k_rs_level<S, 4, RR, EE>)
#define CDG_RS_E(RR)                                  \
    if (EPT <= 2) return CDG_RS_K(RR, 2);             \
    if (EPT <= 4) return CDG_RS_K(RR, 4);             \
    return CDG_RS_K(RR, 8);
    if (R <= 2) { CDG_RS_E(2) }
    if (R <= 3) { CDG_RS_E(3) }
    if (R <= 4) { CDG_RS_E(4) }
    if (R <= 6) { CDG_RS_E(6) }
    CDG_RS_E(8)
#undef CDG_RS_E
#undef CDG_RS_K
}

template <typename S, int K>
const void *rsm_kernel_k(int R) {
    if (R <= 2) return reinterpret_cast<const void *>(&k_rs_multi<S, K, 2>);
    if (R <= 3) return reinterpret_cast<const void *>(&k_rs_multi<S, K, 3>);
    if (R <= 4) return reinterpret_cast<const void *>(&k_rs_multi<S, K, 4>);
    if (R <= 6) return reinterpret_cast<const void *>(&k_rs_multi<S, K, 6>);
    return reinterpret_cast<const void *>(&k_rs_multi<S, K, 8>);
}

template <typename S>
const void *rsm_kernel(int K, int R) {
    if (K <= 2) return rsm_kernel_k<S, 2>(R);
    if (K <= 4) return rsm_kernel_k<S, 4>(R);
    return rsm_kernel_k<S, 8>(R);
}

int round_pow(int x) { return x <= 2 ? 2 : (x <= 4 ? 4 : 8); }
int round_r(int x) { return x <= 2 ? 2 : x <= 3 ? 3 : x <= 4 ? 4 : x <= 6 ? 6 : 8; }

std::atomic<int> g_live_sweepers{0};

}  // namespace

// Live contexts that may run persistent sweeps concurrently (each takes a
// share of the device's resident workgroups, so concurrent sweeps on other
// streams can all be resident at once).
void rs_ctx_count(int delta) { g_live_sweepers += delta; }

// Layout, task list, launch geometry and storage of the row-slice schedule
// for M (once per preconditioner).
int rs_setup(cdg_ctx *ctx, cdg_prec *M) {
    if (M->rs_tasks) return CDG_OK;
    const int n = M->n, prec = M->precision;
    if (M->maxdep + 1 > kRsMaxBlocks)
        return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "stream schedule: a row has more than %d lower or upper blocks",
                        kRsMaxBlocks - 1);
    const int V = prec == CDG_FP64 ? 2 : (prec == CDG_FP32 ? 4 : 8);
    const size_t es = prec == CDG_FP64 ? 8 : (prec == CDG_FP32 ? 4 : 2);
    const size_t accsz = prec == CDG_FP64 ? 8 : 4;
    // rows per slice and threads per task: measured best on the 16^3 P2 Euler
    // system (FP64 15 x ~256, FP32 27 x ~512, BF16 45 x ~512: the larger the
    // slice, the fewer tasks share each parent vector's loads; FP64 slices
    // hold twice the bytes per row)
    int wmax = prec == CDG_FP64 ? 16 : (prec == CDG_FP32 ? 27 : 45);
    if (const char *e = std::getenv("CDG_RS_W")) wmax = std::max(1, std::atoi(e));
    int target = prec == CDG_FP64 ? 256 : 512;
    if (const char *e = std::getenv("CDG_RS_THREADS")) target = std::max(64, std::atoi(e));
    const int w = rs_slice_rows(n, wmax);
    // a task gathers at most (maxdep + 1) parent vectors, at most 8 values per thread
    const int ne_max = (M->maxdep + 1) * n;
    target = std::max(target, (ne_max + 7) / 8);
    const StreamShape sh = stream_shape(w, V, target, kRsMaxThreads);
    if (sh.T == 0) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "stream schedule: no launch shape for n=%d", n);
    const int T = sh.T;
    const int S = n / w;
    const int64_t SL = ((int64_t)w * n + 7) / 8 * 8;
    const int64_t slot = (S * SL + 63) / 64 * 64;
    const int R = round_r(R_needed<int>(w * n, V * T));
    if ((ne_max + T - 1) / T > 8)
        return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "stream schedule: %d threads cannot gather %d parents of n=%d", T,
                        M->maxdep + 1, n);
    const int EPT = round_pow((ne_max + T - 1) / T);
    const void *fn = prec == CDG_FP64 ? rs_kernel<double>(R, EPT)
                   : (prec == CDG_FP32 ? rs_kernel<float>(R, EPT) : rs_kernel<cdg_bf16>(R, EPT));
    if (w > T) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "stream schedule: %d-row slices need more than %d threads", w, T);
    const size_t lds = accsz * ((((size_t)kRsMaxBlocks * n + 3) & ~(size_t)3) + (size_t)V * T + (size_t)T);
    int per_cu = 0;
    CDG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, T, lds));
    if (per_cu < 1) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "stream sweep does not fit a CU");
    // task list: forward levels, then backward levels; rows in level order, slices ascending
    std::vector<int32_t> tasks;
    tasks.reserve((size_t)2 * M->nrows * S * kRsRec);
    auto push = [&](int row, int sl, bool bwd) {
        int32_t rec[kRsRec] = {0};
        rec[0] = row;
        int nb = 0;
        if (!bwd) {
            for (int q = M->h_lo_ptr[row]; q < M->h_lo_ptr[row + 1]; ++q, ++nb) {
                rec[2 + nb] = q;
                rec[2 + kRsMaxBlocks + nb] = M->h_lo_col[q];
            }
        } else {
            rec[2] = M->nlo + M->nup + row;
            rec[2 + kRsMaxBlocks] = row;
            nb = 1;
            for (int u = M->h_up_ptr[row]; u < M->h_up_ptr[row + 1]; ++u, ++nb) {
                rec[2 + nb] = M->nlo + u;
                rec[2 + kRsMaxBlocks + nb] = M->h_up_col[u];
            }
        }
        rec[1] = sl | (nb << 8) | ((bwd ? 1 : 0) << 16);
        tasks.insert(tasks.end(), rec, rec + kRsRec);
    };
    for (int l = 0; l < M->nlev; ++l)
        for (int k = M->lev_row_ptr[l]; k < M->lev_row_ptr[l + 1]; ++k)
            for (int sl = 0; sl < S; ++sl) push(M->h_lev_rows[k], sl, false);
    for (int l = 0; l < M->nblev; ++l)
        for (int k = M->blev_row_ptr[l]; k < M->blev_row_ptr[l + 1]; ++k)
            for (int sl = 0; sl < S; ++sl) push(M->h_blev_rows[k], sl, true);
    const int ntasks = (int)(tasks.size() / kRsRec);
    auto alloc = [&](void **p, size_t bytes) -> int {
        bytes = std::max<size_t>(bytes, 256);
        int e = dev_alloc(ctx, p, bytes);
        if (!e) M->owned.push_back({*p, bytes});
        if (!e) {
            hipError_t he = hipMemsetAsync(*p, 0, bytes, ctx->stream);
            if (he != hipSuccess) e = cdg_fail(ctx, CDG_ERR_HIP, "%s", hipGetErrorString(he));
        }
        return e;
    };
    const int WD = prec == CDG_FP64 ? 2 : 1;
    const size_t vec_words = (size_t)M->nrows * n * WD;
    // one allocation: L slots, then W slots, then Dinv slots
    int e = alloc(&M->Lrs, es * (size_t)(M->nlo + M->nup + M->nrows) * slot);
    if (!e) {
        M->Wrs = static_cast<char *>(M->Lrs) + es * (size_t)M->nlo * slot;
        M->Drs = static_cast<char *>(M->Wrs) + es * (size_t)M->nup * slot;
    }
    if (!e) e = alloc(&M->rs_yb, 8 * vec_words);
    if (!e) e = alloc(&M->rs_y, (prec == CDG_FP64 ? 8 : 4) * (size_t)M->nrows * n);
    if (!e) e = alloc(&M->rs_zb, 8 * vec_words);
    if (!e) e = alloc((void **)&M->d_rs_ctl, sizeof(uint32_t) * kRsCtlWords);
    if (!e) {
        int32_t *dt = nullptr;
        const size_t bytes = sizeof(int32_t) * std::max<size_t>(tasks.size(), 1);
        e = dev_alloc(ctx, (void **)&dt, bytes);
        if (!e) {
            M->owned.push_back({dt, bytes});
            hipError_t he = hipMemcpyAsync(dt, tasks.data(), sizeof(int32_t) * tasks.size(), hipMemcpyHostToDevice,
                                           ctx->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
            if (he != hipSuccess) e = cdg_fail(ctx, CDG_ERR_HIP, "%s", hipGetErrorString(he));
            M->rs_tasks = dt;
        }
    }
    if (!e) {   // the up entry -> row map (the fold's left operand)
        std::vector<int32_t> up_row(std::max(M->nup, 1), 0);
        for (int i = 0; i < M->nrows; ++i)
            for (int u = M->h_up_ptr[i]; u < M->h_up_ptr[i + 1]; ++u) up_row[u] = i;
        int32_t *d = nullptr;
        const size_t bytes = sizeof(int32_t) * up_row.size();
        e = dev_alloc(ctx, (void **)&d, bytes);
        if (!e) {
            M->owned.push_back({d, bytes});
            hipError_t he = hipMemcpyAsync(d, up_row.data(), bytes, hipMemcpyHostToDevice, ctx->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
            if (he != hipSuccess) e = cdg_fail(ctx, CDG_ERR_HIP, "%s", hipGetErrorString(he));
            M->d_up_row = d;
        }
    }
    if (e) return e;
    M->rs_w = w;
    M->rs_S = S;
    M->rs_SL = SL;
    M->rs_slot = slot;
    M->rs_T = T;
    M->rs_dc = sh.dc;
    M->rs_R = R;
    M->rs_EPT = EPT;
    M->rs_lds = lds;
    M->rs_ntasks = ntasks;
    M->rs_per_cu = per_cu;
    M->rs_ready = 0;
    M->rs_spin = 1 << 20;
    // ticket classes: each class's counter serves G / classes workgroups; the
    // first `classes` workgroups (one per class) must be resident together
    M->rs_classes = 64;
    if (const char *ev = std::getenv("CDG_RS_CLASSES")) M->rs_classes = std::atoi(ev);
    M->rs_classes = std::max(1, std::min(M->rs_classes, kRsMaxClass));
    if (const char *ev = std::getenv("CDG_RS_SPIN")) M->rs_spin = std::max(1, std::atoi(ev));
    return CDG_OK;
}

// Resident workgroups for one sweep: the device's capacity shared by the
// live contexts (CDG_RS_GRID caps it).
static int rs_grid(cdg_ctx *ctx, const cdg_prec *M) {
    const int live = std::max(1, g_live_sweepers.load());
    int g = std::max(1, M->rs_per_cu * ctx->num_cus / live);
    if (const char *e = std::getenv("CDG_RS_GRID")) g = std::max(1, std::min(g, std::atoi(e)));
    return std::min(g, M->rs_ntasks);
}

// L and Dinv to the row-slice layout (the fold writes W itself).
int rs_relayout(cdg_ctx *ctx, cdg_prec *M) {
    const int n = M->n, w = M->rs_w, SL = (int)M->rs_SL;
    const dim3 blk(256);
    auto run = [&](const void *src, void *dst, int count) -> int {
        if (count == 0) return CDG_OK;
        const dim3 grid((unsigned)std::min(count, 65535));
        if (M->precision == CDG_FP64)
            hipLaunchKernelGGL((k_to_rs<double, double>), grid, blk, 0, ctx->stream, static_cast<const double *>(src),
                               M->ld, static_cast<double *>(dst), M->rs_slot, count, n, w, SL);
        else if (M->precision == CDG_FP32)
            hipLaunchKernelGGL((k_to_rs<float, float>), grid, blk, 0, ctx->stream, static_cast<const float *>(src),
                               M->ld, static_cast<float *>(dst), M->rs_slot, count, n, w, SL);
        else
            hipLaunchKernelGGL((k_to_rs<float, cdg_bf16>), grid, blk, 0, ctx->stream, static_cast<const float *>(src),
                               M->ld, static_cast<cdg_bf16 *>(dst), M->rs_slot, count, n, w, SL);
        CDG_HIP_TRY(hipGetLastError());
        return CDG_OK;
    };
    const void *Lsrc = M->precision == CDG_BF16 ? static_cast<const void *>(M->L32) : M->L;
    const void *Dsrc = M->precision == CDG_BF16 ? static_cast<const void *>(M->D32) : M->Dinv;
    int e = run(Lsrc, M->Lrs, M->nlo);
    if (!e) e = run(Dsrc, M->Drs, M->nrows);
    return e;
}

namespace {
template <typename S>
int rs_launch_t(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z, uint64_t *yb, uint64_t *zb, uint32_t *ctl,
                uint32_t tag) {
    const int G = rs_grid(ctx, M);
    RsArgs a{};
    a.tasks = M->rs_tasks;
    a.ntasks = M->rs_ntasks;
    a.Sb = M->Lrs;
    a.slot = M->rs_slot;
    a.SL = (int32_t)M->rs_SL;
    a.n = M->n;
    a.w = M->rs_w;
    a.dc = M->rs_dc;
    a.v = v;
    a.z = z;
    a.yb = yb;
    a.zb = zb;
    a.tag = tag;
    a.ctl = ctl;
    a.nclass = std::max(1, std::min(M->rs_classes, G));
    a.spin_limit = M->rs_spin;
    const void *fn = rs_kernel<S>(M->rs_R, M->rs_EPT);
    void *args[] = {&a};
    CDG_HIP_TRY(hipLaunchKernel(fn, dim3(G), dim3(M->rs_T), args, M->rs_lds, ctx->stream));
    return CDG_OK;
}
}  // namespace

// One apply with the row-slice sweep on the given exchange buffers (words of
// nrows*n tagged values each) and control words (rs_ctl_words() uint32: the
// error word and the ticket counters); `tag` must differ from every tag the
// buffers hold (the caller's call counter), and the ticket set of parity
// tag & 1 must be clear (zeroed memory, or cleared by the call with tag - 1).
int launch_ilu_apply_rs_buf(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z, uint64_t *yb, uint64_t *zb,
                            uint32_t *ctl, uint32_t tag) {
    if (M->nrows == 0) return CDG_OK;
    if (M->precision == CDG_FP64) return rs_launch_t<double>(ctx, M, v, z, yb, zb, ctl, tag);
    if (M->precision == CDG_FP32) return rs_launch_t<float>(ctx, M, v, z, yb, zb, ctl, tag);
    return rs_launch_t<cdg_bf16>(ctx, M, v, z, yb, zb, ctl, tag);
}

int rs_ctl_words() { return kRsCtlWords; }

int launch_ilu_apply_rs(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z) {
    if (M->rs_epoch == 0xffffffffu) {   // tags wrap: clear the buffers, restart at 1
        const size_t words = (size_t)M->nrows * M->n * (M->precision == CDG_FP64 ? 2 : 1);
        CDG_HIP_TRY(hipMemsetAsync(M->rs_yb, 0, 8 * words, ctx->stream));
        CDG_HIP_TRY(hipMemsetAsync(M->rs_zb, 0, 8 * words, ctx->stream));
        CDG_HIP_TRY(hipMemsetAsync(M->d_rs_ctl, 0, sizeof(uint32_t) * kRsCtlWords, ctx->stream));
        M->rs_epoch = 0;
    }
    M->rs_epoch += 1;
    return launch_ilu_apply_rs_buf(ctx, M, v, z, static_cast<uint64_t *>(M->rs_yb), static_cast<uint64_t *>(M->rs_zb),
                                   M->d_rs_ctl, M->rs_epoch);
}

// One apply of the row-slice tasks, one launch per wavefront level (forward
// levels, then backward levels), with the forward vector in y (nrows * n
// values of the accumulation type).
int launch_ilu_apply_rsl_buf(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z, void *y) {
    if (M->nrows == 0) return CDG_OK;
    RslArgs a{};
    a.tasks = M->rs_tasks;
    a.Sb = M->Lrs;
    a.slot = M->rs_slot;
    a.SL = (int32_t)M->rs_SL;
    a.n = M->n;
    a.w = M->rs_w;
    a.dc = M->rs_dc;
    a.v = v;
    a.z = z;
    a.y = y;
    const void *fn = M->precision == CDG_FP64 ? rsl_kernel<double>(M->rs_R, M->rs_EPT)
                   : (M->precision == CDG_FP32 ? rsl_kernel<float>(M->rs_R, M->rs_EPT)
                                               : rsl_kernel<cdg_bf16>(M->rs_R, M->rs_EPT));
    const int S = M->rs_S;
    auto run = [&](int lo, int hi) -> int {
        if (hi <= lo) return CDG_OK;
        a.t0 = lo;
        void *args[] = {&a};
        CDG_HIP_TRY(hipLaunchKernel(fn, dim3(hi - lo), dim3(M->rs_T), args, M->rs_lds, ctx->stream));
        return CDG_OK;
    };
    for (int l = 0; l < M->nlev; ++l) {
        int e = run(S * M->lev_row_ptr[l], S * M->lev_row_ptr[l + 1]);
        if (e) return e;
    }
    const int nf = S * M->nrows;
    for (int l = 0; l < M->nblev; ++l) {
        int e = run(nf + S * M->blev_row_ptr[l], nf + S * M->blev_row_ptr[l + 1]);
        if (e) return e;
    }
    return CDG_OK;
}

int launch_ilu_apply_rsl(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z) {
    return launch_ilu_apply_rsl_buf(ctx, M, v, z, M->rs_y);
}

// Right-hand sides per launch of the k-vector apply: as many as keep the
// parents of one task within 48 KB of LDS (1, 2, 4 or 8).
int rs_multi_width(const cdg_prec *M) {
    const size_t accsz = M->precision == CDG_FP64 ? 8 : 4;
    const size_t per = (size_t)(M->maxdep + 1) * M->n * accsz;
    int fit = (int)std::max<size_t>(1, (48 * 1024) / std::max<size_t>(per, 1));
    static const int cap = std::getenv("CDG_RS_MULTI_WIDTH") ? std::max(1, std::atoi(std::getenv("CDG_RS_MULTI_WIDTH"))) : 8;
    fit = std::min(fit, cap);
    return fit >= 8 ? 8 : fit >= 4 ? 4 : fit >= 2 ? 2 : 1;
}

// k applies of the row-slice schedule at once: z_r = M^-1 v_r for r < k
// (v, z: k vectors `stride` apart, device); y: the forward vectors, k
// vectors of nrows * n values of the accumulation type `stride` apart.
// The row-slice storage must be prepared (rs_prepare).  Groups of up to
// rs_multi_width(M) right-hand sides share each block read; a single vector
// runs the single-vector kernel.  row_k (may be null): per row, the number of
// right-hand sides of the row's system (<= k), for a batch of systems joined
// in one matrix; vectors r >= row_k[row] of that row are left alone.
// row_hi >= 0: only the rows row_lo <= i < row_hi (every level's rows are in
// ascending order, so a row range is one task range per level; levels
// without such rows launch nothing).
int launch_ilu_apply_rs_multi(cdg_ctx *ctx, cdg_prec *M, const double *v, double *z, void *y, int64_t stride, int k,
                              const int32_t *row_k, int row_lo, int row_hi) {
    if (M->nrows == 0 || k <= 0) return CDG_OK;
    const int S = M->rs_S, nf = S * M->nrows;
    const bool ranged = row_hi >= 0;
    auto sub = [&](const std::vector<int32_t> &rows, int a, int b, int *lo, int *hi) {
        if (!ranged) { *lo = a; *hi = b; return; }
        *lo = (int)(std::lower_bound(rows.begin() + a, rows.begin() + b, row_lo) - rows.begin());
        *hi = (int)(std::lower_bound(rows.begin() + a, rows.begin() + b, row_hi) - rows.begin());
    };
    const int kw = rs_multi_width(M);
    const size_t vt = M->precision == CDG_FP64 ? 8 : 4;
    for (int r0 = 0; r0 < k; r0 += kw) {
        const int kk = std::min(kw, k - r0);
        const double *vr = v + (size_t)r0 * stride;
        double *zr = z + (size_t)r0 * stride;
        void *yr = static_cast<char *>(y) + vt * (size_t)r0 * stride;
        if (kk == 1 && !row_k && !ranged) {
            int e = launch_ilu_apply_rsl_buf(ctx, M, vr, zr, yr);
            if (e) return e;
            continue;
        }
        const int K = kk <= 2 ? 2 : kk <= 4 ? 4 : 8;
        RsmArgs a{};
        a.tasks = M->rs_tasks;
        a.Sb = M->Lrs;
        a.slot = M->rs_slot;
        a.SL = (int32_t)M->rs_SL;
        a.n = M->n;
        a.w = M->rs_w;
        a.dc = M->rs_dc;
        a.v = vr;
        a.z = zr;
        a.y = yr;
        a.stride = stride;
        a.k = kk;
        a.nbmax = M->maxdep + 1;
        a.row_k = row_k;
        a.r0 = r0;
        const void *fn = M->precision == CDG_FP64 ? rsm_kernel<double>(K, M->rs_R)
                       : (M->precision == CDG_FP32 ? rsm_kernel<float>(K, M->rs_R)
                                                   : rsm_kernel<cdg_bf16>(K, M->rs_R));
        const size_t accsz = M->precision == CDG_FP64 ? 8 : 4;
        const int V = M->precision == CDG_FP64 ? 2 : (M->precision == CDG_FP32 ? 4 : 8);
        const size_t R2 = K < 4 ? K : (V >= 8 ? 2 : 4);   // the kernel's row-sum round width
        const size_t lds = accsz * (std::max((((size_t)K * a.nbmax * M->n + 3) & ~(size_t)3), R2 * V * M->rs_T) +
                                    R2 * M->rs_T);
        auto run = [&](int lo, int hi) -> int {
            if (hi <= lo) return CDG_OK;
            a.t0 = lo;
            void *args[] = {&a};
            CDG_HIP_TRY(hipLaunchKernel(fn, dim3(hi - lo), dim3(M->rs_T), args, lds, ctx->stream));
            return CDG_OK;
        };
        for (int l = 0; l < M->nlev; ++l) {
            int lo, hi;
            sub(M->h_lev_rows, M->lev_row_ptr[l], M->lev_row_ptr[l + 1], &lo, &hi);
            int e = run(S * lo, S * hi);
            if (e) return e;
        }
        for (int l = 0; l < M->nblev; ++l) {
            int lo, hi;
            sub(M->h_blev_rows, M->blev_row_ptr[l], M->blev_row_ptr[l + 1], &lo, &hi);
            int e = run(nf + S * lo, nf + S * hi);
            if (e) return e;
        }
    }
    return CDG_OK;
}
