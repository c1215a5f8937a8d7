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
#include <type_traits>

namespace {

constexpr int kMaxThreads = 1024;

typedef double nat_f64x2 __attribute__((ext_vector_type(2)));

// ---- 16-byte vector loads ---------------------------------------------------
template <typename S, typename Acc> struct Ld16;

template <> struct Ld16<double, double> {
    static constexpr int V = 2;
    __device__ static __forceinline__ void load(const double *p, double (&o)[2]) {
        const nat_f64x2 t = __builtin_nontemporal_load(reinterpret_cast<const nat_f64x2 *>(p));
        o[0] = t.x; o[1] = t.y;
    }
};
typedef float nat_f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t nat_u32x4 __attribute__((ext_vector_type(4)));
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

// acc[v] += sum over this thread's elements of blk(e) * xs(col(e)), from
// walk step m0 on; U loads of 16 bytes in flight per thread.
template <typename S, typename Acc, int U = 4>
__device__ __forceinline__ void stream_block(const S *__restrict__ blk, const Acc *__restrict__ xs, int nn,
                                             const Walk<Ld16<S, Acc>::V> &w, Acc *acc, int m0 = 0) {
    constexpr int V = Ld16<S, Acc>::V;
    const int mfull = nn / w.step;           // steps where every element is in range
    int m = m0;
    for (; m + U <= mfull; m += U) {
        Acc a[U][V];
#pragma unroll
        for (int u = 0; u < U; ++u) Ld16<S, Acc>::load(blk + w.e0 + (m + u) * w.step, a[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
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
template <int U>
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
            stream_block<double, double, U>(blocks + (size_t)(cb + k) * ld, xs + k * n, nn, w, acc);
    }
    park<double, V>(red, tid, acc);
    __syncthreads();
    const int parts = V * T / n;
    for (int r = tid; r < n; r += T) y[(size_t)row * n + r] = row_sum(red, r, n, parts);
}

// ============================================================================
// K3: ILU(0) apply (precon_ILU0.f90:286-411), level-synchronous.
//   forward,  level l:  z_i = v_i - sum_j L_ij z_j            (rows of level l)
//   backward, level l:  t_i = z_i - sum_k U_ik z_k   (stage A)
//                       z_i = Dinv_i t_i             (stage B)
// Rows of one wavefront level are independent, so a level is one launch.  Its
// workgroups own (row, column slice) pairs: a slice of columns of a
// column-major block is one contiguous range, streamed with the same
// fixed-row 16-byte walk as the matvec.  Each workgroup writes the partial
// n-vector of its slice; vectors are kept in the implicit form
//     u = base + sign1 * sum_s part1[s] + sign2 * sum_s part2[s]
// and rebuilt by their consumers (only the consumer's column slice) in a
// fixed order, so results are deterministic.  A last elementwise pass
// materializes z.  The partials are stored in the accumulation type (FP32 for
// the FP32 and BF16 factors: a slice's sum is an FP32 value, so storing it as
// FP32 loses nothing), and widened to FP64 where they are summed.
// ============================================================================
struct VecView {
    const double *base;       // may be null (0)
    const void *p1, *p2;      // partial arrays [nrows][S][n] of the accumulation type;
                              // p1 never null, p2 may be
    double s1, s2;            // signs of the partial sums
};

// Element c of row `row` of a view is gathered in two steps: view_load issues
// the loads (the base and NB slices of p1 and, with P2, of p2, all from
// clamped in-bounds addresses and without branches), view_sum adds them.
// The partials are summed in slice order s = 0, 1, ..., S-1.  Splitting the
// two lets a kernel put other loads (the first block chunks) in between: the
// load counter is in order, so those stay in flight while the sums wait.
template <int NB> struct ViewLd {
    double b;
    double q1[NB], q2[NB];
};

template <int NB, bool P2, typename PT>
__device__ __forceinline__ ViewLd<NB> view_load(const VecView &u, int row, int c, int n, int S, int s0) {
    ViewLd<NB> r;
    const PT *p1 = static_cast<const PT *>(u.p1), *p2 = static_cast<const PT *>(u.p2);
    const size_t rb = (size_t)row * S * n + c;
    // without a base, an in-bounds stand-in inside p1's bytes (value unused)
    r.b = *(u.base ? u.base + (size_t)row * n + c : reinterpret_cast<const double *>(p1) + rb * sizeof(PT) / 8);
#pragma unroll
    for (int i = 0; i < NB; ++i) r.q1[i] = static_cast<double>(p1[rb + (size_t)min(s0 + i, S - 1) * n]);
#pragma unroll
    for (int i = 0; i < NB; ++i) r.q2[i] = P2 ? static_cast<double>(p2[rb + (size_t)min(s0 + i, S - 1) * n]) : 0.0;
    return r;
}

template <int NB>
__device__ __forceinline__ void view_add(const ViewLd<NB> &r, int S, int s0, double &t1, double &t2) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
        if (s0 + i < S) {
            t1 += r.q1[i];
            t2 += r.q2[i];
        }
}

// u(row, c) from the loads of its first NB slices (the rest loaded here)
template <int NB, bool P2, typename PT>
__device__ __forceinline__ double view_sum(const VecView &u, const ViewLd<NB> &r, int row, int c, int n, int S) {
    double t1 = 0.0, t2 = 0.0;
    view_add<NB>(r, S, 0, t1, t2);
    for (int s0 = NB; s0 < S; s0 += NB) view_add<NB>(view_load<NB, P2, PT>(u, row, c, n, S, s0), S, s0, t1, t2);
    double x = u.base ? r.b : 0.0;
    x += u.s1 * t1;
    if (P2) x += u.s2 * t2;
    return x;
}

template <int NB, bool P2, typename PT>
__device__ __forceinline__ double view_at(const VecView &u, int row, int c, int n, int S) {
    return view_sum<NB, P2, PT>(u, view_load<NB, P2, PT>(u, row, c, n, S, 0), row, c, n, S);
}

struct KSliceArgs {
    const void *B;            // block storage (L, U, or Dinv)
    int64_t ld;               // elements per block slot
    const int32_t *rows;      // rows of this level
    const int32_t *ptr;       // per-row entry range (lower or upper lists); null: the row's own block
    const int32_t *idx;       // block slot of entry d (null: slot = d)
    const int32_t *col;       // parent row of entry d
    VecView x;                // vector the parents' slices are read from
    void *part;               // output partials [nrows][S][n], accumulation type
    int n, S, w, maxdep, dc;  // block size, slices, slice width (columns), max entries, walk step
};

template <typename S, typename Acc, int U, int NB, bool P2>
__global__ void __launch_bounds__(kMaxThreads) k_kslice(KSliceArgs a) {
    constexpr int V = Ld16<S, Acc>::V;
    extern __shared__ unsigned char smem_raw[];
    const int T = blockDim.x, tid = threadIdx.x, n = a.n;
    Acc *xs = reinterpret_cast<Acc *>(smem_raw);             // maxdep * w
    Acc *red = xs + (size_t)a.maxdep * a.w;                   // V*T
    const int row = a.rows[blockIdx.x / a.S];
    const int sl = blockIdx.x - (blockIdx.x / a.S) * a.S;
    const int c0 = sl * a.w;
    const int c1 = min(n, c0 + a.w);
    const int wc = c1 - c0;
    int d0 = 0, d1 = 1;
    if (a.ptr) { d0 = a.ptr[row]; d1 = a.ptr[row + 1]; }
    const int nd = d1 - d0;
    const Walk<V> wk = make_walk<V>(tid, T, n, a.dc);
    const S *Bp = static_cast<const S *>(a.B);
    auto slot_of = [&](int k) { return a.ptr ? (a.idx ? a.idx[d0 + k] : d0 + k) : row; };
    // the slice is the flat range [c0*n, c1*n) of each block; column index
    // inside the slice = global column - c0, so xs is read at (col - c0).
    // The first block's first U walk steps are loaded before the gather.
    // Order of the loads: this thread's first gather element, then the
    // first block's first U walk steps, then (after the sums) the rest.
    // Both go out unconditionally from in-bounds addresses (threads beyond
    // the gather read a valid element, a row without entries reads the
    // partial array), so no branch separates them from the sums: a branch
    // would make the compiler's load-counter waits cover the block loads too.
    const int ne = nd * wc;
    const bool g0 = tid < ne;
    const int kg = g0 ? tid / wc : 0, cg = g0 ? tid - kg * wc : 0;
    const int pg = a.ptr && nd > 0 ? a.col[d0 + kg] : row;
    const ViewLd<NB> r0 = view_load<NB, P2, Acc>(a.x, pg, c0 + cg, n, a.S, 0);
    const int m0 = nd > 0 ? min(U, wc * n / wk.step) : 0;
    const S *b0 = nd > 0 ? Bp + (size_t)slot_of(0) * a.ld + (size_t)c0 * n : reinterpret_cast<const S *>(a.x.p1);
    Acc a0[U][V];
#pragma unroll
    for (int u = 0; u < U; ++u) Ld16<S, Acc>::load(b0 + (u < m0 ? wk.e0 + u * wk.step : 0), a0[u]);
    if (g0) xs[kg * a.w + cg] = static_cast<Acc>(view_sum<NB, P2, Acc>(a.x, r0, pg, c0 + cg, n, a.S));
    for (int e = tid + T; e < ne; e += T) {
        const int k = e / wc, c = e - k * wc;
        const int parent = a.ptr ? a.col[d0 + k] : row;
        xs[k * a.w + c] = static_cast<Acc>(view_at<NB, P2, Acc>(a.x, parent, c0 + c, n, a.S));
    }
    __syncthreads();
    Acc acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0;
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (u < m0)
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] += a0[u][v] * xs[wk.c0[v] + u * wk.dc];
    for (int k = 0; k < nd; ++k)
        stream_block<S, Acc, U>(Bp + (size_t)slot_of(k) * a.ld + (size_t)c0 * n, xs + k * a.w, wc * n, wk, acc,
                                k == 0 ? m0 : 0);
#pragma unroll
    for (int v = 0; v < V; ++v) red[V * tid + v] = acc[v];
    __syncthreads();
    const int parts = V * T / n;
    Acc *out = static_cast<Acc *>(a.part) + ((size_t)row * a.S + sl) * n;
    for (int r = tid; r < n; r += T) out[r] = row_sum(red, r, n, parts);
}

// z_i = Dinv_i t_i for every row i, the folded schedule's last stage (t in
// its partial-sum view).  One workgroup per row streams the whole block and
// writes z_i itself; the block's first U walk steps are loaded before t_i is
// gathered into LDS, so the two latencies overlap.
template <typename S, typename Acc, int U, int NB>
__global__ void __launch_bounds__(kMaxThreads) k_dinv_z(const S *__restrict__ D, int64_t ld, VecView t, int Sp,
                                                        int n, int dc, double *__restrict__ z) {
    constexpr int V = Ld16<S, Acc>::V;
    extern __shared__ unsigned char smem_raw[];
    const int T = blockDim.x, tid = threadIdx.x, row = blockIdx.x;
    Acc *xs = reinterpret_cast<Acc *>(smem_raw);   // n (n + V*T Acc in all, 16-byte aligned pieces)
    Acc *red = xs + ((n + 3) & ~3);                // V*T
    const S *blk = D + (size_t)row * ld;
    const int nn = n * n;
    const Walk<V> w = make_walk<V>(tid, T, n, dc);
    const ViewLd<NB> r0 = view_load<NB, true, Acc>(t, row, min(tid, n - 1), n, Sp, 0);
    const int m0 = min(U, nn / w.step);
    Acc a0[U][V];
#pragma unroll
    for (int u = 0; u < U; ++u) Ld16<S, Acc>::load(blk + (u < m0 ? w.e0 + u * w.step : 0), a0[u]);
    if (tid < n) xs[tid] = static_cast<Acc>(view_sum<NB, true, Acc>(t, r0, row, tid, n, Sp));
    for (int c = tid + T; c < n; c += T) xs[c] = static_cast<Acc>(view_at<NB, true, Acc>(t, row, c, n, Sp));
    __syncthreads();
    Acc acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0;
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (u < m0)
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] += a0[u][v] * xs[w.c0[v] + u * w.dc];
    stream_block<S, Acc, U>(blk, xs, nn, w, acc, m0);
#pragma unroll
    for (int v = 0; v < V; ++v) red[V * tid + v] = acc[v];
    __syncthreads();
    const int parts = V * T / n;
    for (int r = tid; r < n; r += T) z[(size_t)row * n + r] = static_cast<double>(row_sum(red, r, n, parts));
}

// z = sum_s part[s]  (materializes the backward result)
template <typename PT>
__global__ void k_sum_parts(const PT *__restrict__ part, int S, int n, int64_t len, double *__restrict__ z) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < len; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = e / n, c = e - row * n;
        double t = 0.0;
        for (int s = 0; s < S; ++s) t += part[(row * S + s) * n + c];
        z[e] = t;
    }
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
StreamShape stream_shape(int n, int V, int target, int tmax) {
    int g = n;
    {   // g = n / gcd(n, V): T must be a multiple of g for V*T % n == 0
        int a = n, b = V;
        while (b) { int t = a % b; a = b; b = t; }
        g = n / a;
    }
    StreamShape best;
    double best_score = -1e30;
    tmax = std::min(tmax, kMaxThreads);
    for (int T = g; T <= tmax; T += g) {
        if (T < 64 && T + g <= tmax) continue;
        const int waves = (T + 63) / 64;
        const double util = (double)T / (64.0 * waves);
        const double score = util - 0.03 * std::abs(T - target) / 64.0;
        if (score > best_score) { best_score = score; best.T = T; }
    }
    if (best.T == 0 && g <= tmax) best.T = g;
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
    if (A->mv_unroll == 16)
        hipLaunchKernelGGL(k_mv<16>, dim3(A->nrows), dim3(s.T), lds, ctx->stream, A->blocks, A->ld, A->d_row_ptr,
                           A->d_col, x, y, A->n, s.dc, chunk);
    else if (A->mv_unroll == 8)
        hipLaunchKernelGGL(k_mv<8>, dim3(A->nrows), dim3(s.T), lds, ctx->stream, A->blocks, A->ld, A->d_row_ptr,
                           A->d_col, x, y, A->n, s.dc, chunk);
    else
        hipLaunchKernelGGL(k_mv<4>, dim3(A->nrows), dim3(s.T), lds, ctx->stream, A->blocks, A->ld, A->d_row_ptr,
                           A->d_col, x, y, A->n, s.dc, chunk);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

namespace {

template <typename S, typename Acc>
int launch_kslice(cdg_ctx *ctx, const cdg_prec *M, KSliceArgs a, int nrows_lvl) {
    if (nrows_lvl == 0) return CDG_OK;
    constexpr int V = Ld16<S, Acc>::V;
    const StreamShape &sh = M->shape[M->precision];
    a.n = M->n;
    a.S = M->slices;
    a.w = M->slice_w;
    a.maxdep = M->maxdep;
    a.dc = sh.dc;
    const size_t lds = sizeof(Acc) * ((size_t)a.maxdep * a.w + (size_t)V * sh.T);
    const dim3 grid(nrows_lvl * M->slices), block(sh.T);
    // gather batch = the slice count up to 8 (no duplicate loads), batches of 8 beyond
    const bool p2 = a.x.p2 != nullptr;
    const int nb = std::min(a.S, 8);
#define CDG_KSLICE(UU, NBB)                                                                                   \
    do {                                                                                                      \
        if (p2) hipLaunchKernelGGL((k_kslice<S, Acc, UU, NBB, true>), grid, block, lds, ctx->stream, a);      \
        else hipLaunchKernelGGL((k_kslice<S, Acc, UU, NBB, false>), grid, block, lds, ctx->stream, a);        \
    } while (0)
#define CDG_KSLICE_NB(UU)                                                                                     \
    switch (nb) {                                                                                             \
    case 1: CDG_KSLICE(UU, 1); break;                                                                         \
    case 2: CDG_KSLICE(UU, 2); break;                                                                         \
    case 3: CDG_KSLICE(UU, 3); break;                                                                         \
    case 4: CDG_KSLICE(UU, 4); break;                                                                         \
    case 5: CDG_KSLICE(UU, 5); break;                                                                         \
    case 6: CDG_KSLICE(UU, 6); break;                                                                         \
    case 7: CDG_KSLICE(UU, 7); break;                                                                         \
    default: CDG_KSLICE(UU, 8); break;                                                                        \
    }
    if (M->unroll == 8) {
        CDG_KSLICE_NB(8)
    } else {
        CDG_KSLICE_NB(4)
    }
#undef CDG_KSLICE_NB
#undef CDG_KSLICE
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

template <typename S, typename Acc>
int sweep(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    const void *U = M->precision == CDG_FP64 ? static_cast<const void *>(A->blocks) : M->U;
    const int64_t ldu = M->precision == CDG_FP64 ? A->ld : M->ld;
    // partial arrays in the accumulation type (the storage holds 3 part_len doubles)
    Acc *P = reinterpret_cast<Acc *>(M->d_part), *QA = P + M->part_len, *PB = P + 2 * M->part_len;
    const VecView zf{v, P, nullptr, -1.0, 0.0};      // forward result  v - sum P
    const VecView zb{nullptr, PB, nullptr, 1.0, 0.0};  // backward result sum PB
    const VecView t{v, P, QA, -1.0, -1.0};            // stage A result v - sum P - sum QA
    for (int l = 0; l < M->nlev; ++l) {
        const int r0 = M->lev_row_ptr[l], r1 = M->lev_row_ptr[l + 1];
        KSliceArgs f{};
        f.B = M->L; f.ld = M->ld; f.rows = M->d_lev_rows + r0;
        f.ptr = M->d_lo_ptr; f.idx = nullptr; f.col = M->d_lo_col;
        f.x = zf; f.part = P;
        int e = launch_kslice<S, Acc>(ctx, M, f, r1 - r0);
        if (e) return e;
    }
    if (!std::is_same<S, cdg_bf16>::value && M->w_ready && M->apply_mode != CDG_APPLY_LEVELS) {
        // backward with W_ik = U_ik Dinv_k: t_i = y_i - sum_k W_ik t_k, one stage
        // per level, then z = Dinv t for all rows at once
        for (int l = 0; l < M->nblev; ++l) {
            const int r0 = M->blev_row_ptr[l], r1 = M->blev_row_ptr[l + 1];
            KSliceArgs u{};
            u.B = M->W; u.ld = M->ld; u.rows = M->d_blev_rows + r0;
            u.ptr = M->d_up_ptr; u.idx = nullptr; u.col = M->d_up_col;
            u.x = t; u.part = QA;
            int e = launch_kslice<S, Acc>(ctx, M, u, r1 - r0);
            if (e) return e;
        }
        constexpr int V = Ld16<S, Acc>::V;
        const StreamShape &sh = M->shape[M->precision];
        const size_t lds = sizeof(Acc) * ((size_t)((M->n + 3) & ~3) + (size_t)V * sh.T);
        using SD = typename std::conditional<std::is_same<S, cdg_bf16>::value, float, S>::type;   // no BF16 W
        const SD *Dv = static_cast<const SD *>(M->Dinv);
#define CDG_DINVZ(UU, NBB) hipLaunchKernelGGL((k_dinv_z<SD, Acc, UU, NBB>), dim3(M->nrows), dim3(sh.T), lds, \
                                              ctx->stream, Dv, M->ld, t, M->slices, M->n, sh.dc, z)
#define CDG_DINVZ_NB(UU)                                                                                      \
        switch (std::min(M->slices, 8)) {                                                                     \
        case 1: CDG_DINVZ(UU, 1); break;                                                                      \
        case 2: CDG_DINVZ(UU, 2); break;                                                                      \
        case 3: CDG_DINVZ(UU, 3); break;                                                                      \
        case 4: CDG_DINVZ(UU, 4); break;                                                                      \
        case 5: CDG_DINVZ(UU, 5); break;                                                                      \
        case 6: CDG_DINVZ(UU, 6); break;                                                                      \
        case 7: CDG_DINVZ(UU, 7); break;                                                                      \
        default: CDG_DINVZ(UU, 8); break;                                                                     \
        }
        if (M->unroll == 8) {
            CDG_DINVZ_NB(8)
        } else {
            CDG_DINVZ_NB(4)
        }
#undef CDG_DINVZ_NB
#undef CDG_DINVZ
        CDG_HIP_TRY(hipGetLastError());
        return CDG_OK;
    } else {   // U and Dinv stages per level (BF16: no W)
        for (int l = 0; l < M->nblev; ++l) {
            const int r0 = M->blev_row_ptr[l], r1 = M->blev_row_ptr[l + 1];
            KSliceArgs u{};
            u.B = U; u.ld = ldu; u.rows = M->d_blev_rows + r0;
            u.ptr = M->d_up_ptr; u.idx = M->d_up_idx; u.col = M->d_up_col;
            u.x = zb; u.part = QA;
            int e = launch_kslice<S, Acc>(ctx, M, u, r1 - r0);
            if (e) return e;
            KSliceArgs d{};
            d.B = M->Dinv; d.ld = M->ld; d.rows = M->d_blev_rows + r0;
            d.ptr = nullptr; d.idx = nullptr; d.col = nullptr;
            d.x = t; d.part = PB;
            e = launch_kslice<S, Acc>(ctx, M, d, r1 - r0);
            if (e) return e;
        }
    }
    const int64_t len = (int64_t)M->nrows * M->n;
    hipLaunchKernelGGL(k_sum_parts<Acc>, dim3((unsigned)std::min<int64_t>((len + 255) / 256, 4096)), dim3(256), 0,
                       ctx->stream, PB, M->slices, M->n, len, z);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

}  // namespace

int launch_ilu_apply(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    if (M->nrows == 0) return CDG_OK;
    if (M->precision == CDG_FP64) return sweep<double, double>(ctx, M, A, v, z);
    if (M->precision == CDG_FP32) return sweep<float, float>(ctx, M, A, v, z);
    return sweep<cdg_bf16, float>(ctx, M, A, v, z);
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
