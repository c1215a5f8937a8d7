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

// ---- 16-byte vector loads ---------------------------------------------------
template <typename S, typename Acc> struct Ld16;

template <> struct Ld16<double, double> {
    static constexpr int V = 2;
    __device__ static __forceinline__ void load(const double *p, double (&o)[2]) {
        const nat_f64x2 t = __builtin_nontemporal_load(reinterpret_cast<const nat_f64x2 *>(p));
        o[0] = t.x; o[1] = t.y;
    }
};
template <typename S, typename Acc> __device__ __forceinline__ Acc to_acc(S v) { return static_cast<Acc>(v); }
template <> __device__ __forceinline__ float to_acc<cdg_bf16, float>(cdg_bf16 v) {
    return __uint_as_float(static_cast<uint32_t>(v) << 16);
}

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
// K3: ILU(0) apply (precon_ILU0.f90:286-411), level-synchronous.
//   forward,  level l:  z_i = v_i - sum_j L_ij z_j            (rows of level l)
//   backward, level l:  t_i = z_i - sum_k U_ik z_k   (stage A)
//                       z_i = Dinv_i t_i             (stage B)
// Rows of one wavefront level are independent, so a level is one launch whose
// workgroups each own one (row, output slice) pair: the slice's rows of every
// block are read column by column (a column's slice is contiguous), every
// thread keeps one fixed output row and a strided set of columns, and the
// partial sums meet in LDS in a fixed order.  No inter-workgroup waiting.
// ============================================================================
struct SliceArgs {
    const void *B;            // block storage (L, U, or Dinv)
    int64_t ld;               // elements per block slot
    const int32_t *rows;      // rows of this level
    const int32_t *ptr;       // per-row block range (lower or upper lists); null: the row's own block
    const int32_t *idx;       // block slot of entry d (null: slot = d)
    const int32_t *col;       // parent row of entry d
    const double *x;          // vector the parents are read from
    const double *base;       // result = base_i - sum (null: result = sum)
    double *out;              // result vector
    int n, S, len;            // block size, slices per row, slice length
    int maxdep;
};

template <typename S, typename Acc>
__global__ void __launch_bounds__(kMaxThreads) k_slice_gemv(SliceArgs a) {
    extern __shared__ unsigned char smem_raw[];
    const int T = blockDim.x, tid = threadIdx.x, n = a.n, len = a.len;
    Acc *xs = reinterpret_cast<Acc *>(smem_raw);            // maxdep * n
    Acc *red = xs + (size_t)a.maxdep * n;                    // T
    const int row = a.rows[blockIdx.x / a.S];
    const int slice = blockIdx.x - (blockIdx.x / a.S) * a.S;
    const int r0 = slice * len;
    const int r1 = min(n, r0 + len);
    int d0 = 0, d1 = 1;
    if (a.ptr) { d0 = a.ptr[row]; d1 = a.ptr[row + 1]; }
    const int nd = d1 - d0;
    for (int e = tid; e < nd * n; e += T) {
        const int k = e / n, c = e - k * n;
        const int parent = a.ptr ? a.col[d0 + k] : row;
        xs[e] = static_cast<Acc>(a.x[(size_t)parent * n + c]);
    }
    __syncthreads();
    const int rr = tid % len, cg = tid / len, ncg = T / len;
    const S *Bp = static_cast<const S *>(a.B);
    Acc acc = 0;
    if (cg < ncg && r0 + rr < r1) {
        for (int k = 0; k < nd; ++k) {
            const int slot = a.ptr ? (a.idx ? a.idx[d0 + k] : d0 + k) : row;
            const S *blk = Bp + (size_t)slot * a.ld + r0 + rr;
            const Acc *xk = xs + k * n;
            int c = cg;
            for (; c + 3 * ncg < n; c += 4 * ncg) {
                const Acc b0 = to_acc<S, Acc>(blk[(size_t)c * n]);
                const Acc b1 = to_acc<S, Acc>(blk[(size_t)(c + ncg) * n]);
                const Acc b2 = to_acc<S, Acc>(blk[(size_t)(c + 2 * ncg) * n]);
                const Acc b3 = to_acc<S, Acc>(blk[(size_t)(c + 3 * ncg) * n]);
                acc += b0 * xk[c];
                acc += b1 * xk[c + ncg];
                acc += b2 * xk[c + 2 * ncg];
                acc += b3 * xk[c + 3 * ncg];
            }
            for (; c < n; c += ncg) acc += to_acc<S, Acc>(blk[(size_t)c * n]) * xk[c];
        }
    }
    red[tid] = acc;
    __syncthreads();
    if (tid < r1 - r0) {
        Acc s = 0;
        for (int g = 0; g < ncg; ++g) s += red[g * len + tid];
        const size_t o = (size_t)row * n + r0 + tid;
        a.out[o] = a.base ? a.base[o] - static_cast<double>(s) : static_cast<double>(s);
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

namespace {

template <typename S, typename Acc>
int launch_slice(cdg_ctx *ctx, const cdg_prec *M, SliceArgs a, int nrows_lvl) {
    if (nrows_lvl == 0) return CDG_OK;
    a.n = M->n;
    a.S = M->slices;
    a.len = M->slice_len;
    a.maxdep = M->maxdep;
    const int T = M->slice_threads;
    const size_t lds = sizeof(Acc) * ((size_t)a.maxdep * M->n + T);
    hipLaunchKernelGGL((k_slice_gemv<S, Acc>), dim3(nrows_lvl * M->slices), dim3(T), lds, ctx->stream, a);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

template <typename S, typename Acc>
int sweep(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    const void *U = M->precision == CDG_FP64 ? static_cast<const void *>(A->blocks) : M->U;
    const int64_t ldu = M->precision == CDG_FP64 ? A->ld : M->ld;
    for (int l = 0; l < M->nlev; ++l) {
        const int r0 = M->lev_row_ptr[l], r1 = M->lev_row_ptr[l + 1];
        SliceArgs f{};
        f.B = M->L; f.ld = M->ld; f.rows = M->d_lev_rows + r0;
        f.ptr = M->d_lo_ptr; f.idx = nullptr; f.col = M->d_lo_col;
        f.x = z; f.base = v; f.out = z;
        int e = launch_slice<S, Acc>(ctx, M, f, r1 - r0);
        if (e) return e;
    }
    for (int l = 0; l < M->nblev; ++l) {
        const int r0 = M->blev_row_ptr[l], r1 = M->blev_row_ptr[l + 1];
        SliceArgs u{};
        u.B = U; u.ld = ldu; u.rows = M->d_blev_rows + r0;
        u.ptr = M->d_up_ptr; u.idx = M->d_up_idx; u.col = M->d_up_col;
        u.x = z; u.base = z; u.out = M->d_t;
        int e = launch_slice<S, Acc>(ctx, M, u, r1 - r0);
        if (e) return e;
        SliceArgs d{};
        d.B = M->Dinv; d.ld = M->ld; d.rows = M->d_blev_rows + r0;
        d.ptr = nullptr; d.idx = nullptr; d.col = nullptr;
        d.x = M->d_t; d.base = nullptr; d.out = z;
        e = launch_slice<S, Acc>(ctx, M, d, r1 - r0);
        if (e) return e;
    }
    return CDG_OK;
}

}  // namespace

int launch_ilu_apply(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const double *v, double *z) {
    if (M->nrows == 0) return CDG_OK;
    if (v == z) return cdg_fail(ctx, CDG_ERR_ARG, "ILU apply needs distinct input and output vectors");
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
