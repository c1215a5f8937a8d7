// Kernels for k right-hand sides of one system that share its matrix and its
// ILU(0) factor: every block is read from HBM once per call for all k
// vectors, so a block's bytes are amortized over k products (arithmetic
// intensity 0.25 k flop/B in FP64 instead of 0.25).
//
//  * k_mv_multi:   Y_r = A X_r, r < k      (chidg_mv, operator_chidg_mv.f90:200-374,
//                                            once per right-hand side in the reference)
//  * k_rs_multi:   the row-slice ILU(0) apply (precon_ILU0.f90:286-411) of
//                  k vectors: forward y_r = v_r - L y_r, backward
//                  z_r = Dinv y_r - W z_r with W = Dinv U, one launch per
//                  wavefront level, tasks (block row, row slice) as in the
//                  single-vector row-slice schedule (k_rsapply.hip).
//
// Vectors of the k right-hand sides lie `stride` elements apart.  Each kernel
// sums in a fixed order that does not depend on k or on timing, and each
// right-hand side's arithmetic is exactly the single-vector kernel's
// (the same products in the same order), so a right-hand side gets the same
// bits alone and among others.
#include "cdg_internal.h"

#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace {

constexpr int kMvMaxThreads = 1024;
typedef double mk_f64x2 __attribute__((ext_vector_type(2)));

// y_r = A x_r for r < k (K >= k right-hand sides per launch).  One
// workgroup per block row; thread t reads 2 doubles at flat offsets
// 2t + 2Tm of each column-major block (2T a multiple of n: the thread's rows
// are fixed, its column advances by dc = 2T/n per step), parent vectors of a
// chunk of blocks staged in LDS for all k right-hand sides.
template <int K, int U>
__global__ void __launch_bounds__(kMvMaxThreads)
k_mv_multi(const double *__restrict__ blocks, int64_t ld, const int32_t *__restrict__ row_ptr,
           const int32_t *__restrict__ col, const double *__restrict__ X, double *__restrict__ Y, int64_t stride,
           int k, int n, int dc, int chunk, const int32_t *__restrict__ row_k, int r0, int row0) {
    extern __shared__ double smem[];
    constexpr int V = 2;
    const int T = blockDim.x, tid = threadIdx.x, row = row0 + (int)blockIdx.x;
    if (row_k) k = min(max(row_k[row] - r0, 0), K);   // the row's system has row_k[row] right-hand sides
    if (k == 0) return;
    const int nn = n * n;
    double *xs = smem;                            // chunk * K * n: [block][rhs][c]
    double *red = smem + (size_t)chunk * K * n;   // V * T
    const int e0 = V * tid, step = V * T;
    const int c0 = e0 / n, c1 = (e0 + 1) / n;
    const bool odd = n & 1;                        // even n: both values of a pair in one column
    double acc[K][V];
#pragma unroll
    for (int r = 0; r < K; ++r) acc[r][0] = acc[r][1] = 0.0;
    const int b0 = row_ptr[row], b1 = row_ptr[row + 1];
    const int mfull = nn / step;
    for (int cb = b0; cb < b1; cb += chunk) {
        const int nb = min(chunk, b1 - cb);
        if (cb != b0) __syncthreads();
        for (int idx = tid; idx < nb * n; idx += T) {   // each parent value's k loads together
            const int kb = idx / n, c = idx - kb * n;
            const size_t gb = (size_t)col[cb + kb] * n + c;
            double xv[K];
#pragma unroll
            for (int r = 0; r < K; ++r) xv[r] = r < k ? X[(size_t)r * stride + gb] : 0.0;
#pragma unroll
            for (int r = 0; r < K; ++r) xs[(size_t)kb * K * n + r * n + c] = xv[r];
        }
        __syncthreads();
        for (int kb = 0; kb < nb; ++kb) {
            const double *blk = blocks + (size_t)(cb + kb) * ld;
            const double *xk = xs + (size_t)kb * K * n;
            auto fma_step = [&](const mk_f64x2 a, int m) {
#pragma unroll
                for (int r = 0; r < K; ++r) {
                    const double xa = xk[r * n + c0 + m * dc];
                    const double xb = odd ? xk[r * n + c1 + m * dc] : xa;
                    acc[r][0] += a.x * xa;
                    acc[r][1] += a.y * xb;
                }
            };
            int m = 0;
            for (; m + U <= mfull; m += U) {
                mk_f64x2 a[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    a[u] = __builtin_nontemporal_load(reinterpret_cast<const mk_f64x2 *>(blk + e0 + (m + u) * step));
#pragma unroll
                for (int u = 0; u < U; ++u) fma_step(a[u], m + u);
            }
            for (; m < mfull; ++m)
                fma_step(__builtin_nontemporal_load(reinterpret_cast<const mk_f64x2 *>(blk + e0 + m * step)), m);
            const int e = e0 + mfull * step;
            if (e < nn) {   // tail: the 16-byte load stays inside the padded slot
                const mk_f64x2 a = __builtin_nontemporal_load(reinterpret_cast<const mk_f64x2 *>(blk + e));
#pragma unroll
                for (int r = 0; r < K; ++r) {
                    if (e < nn) acc[r][0] += a.x * xk[r * n + c0 + mfull * dc];
                    if (e + 1 < nn) acc[r][1] += a.y * xk[r * n + c1 + mfull * dc];
                }
            }
        }
    }
    // row sums per right-hand side, in the single-vector kernel's order
    const int parts = V * T / n;
#pragma unroll
    for (int r = 0; r < K; ++r) {
        if (r >= k) break;
        __syncthreads();
        red[V * tid] = acc[r][0];
        red[V * tid + 1] = acc[r][1];
        __syncthreads();
        for (int i = tid; i < n; i += T) {
            double s = 0.0;
            for (int q = 0; q < parts; ++q) s += red[i + n * q];
            Y[(size_t)r * stride + (size_t)row * n + i] = s;
        }
    }
}

}  // namespace

// Y_r = A X_r for r < k (X, Y: k vectors `stride` elements apart, device).
// row_k (may be null): per row, the right-hand sides of the row's system
// (<= k), for a batch of systems joined in one matrix.  row_hi >= 0: only
// the rows row_lo <= i < row_hi.
int launch_mv_multi(cdg_ctx *ctx, const cdg_mat *A, const double *X, double *Y, int64_t stride, int k,
                    const int32_t *row_k, int row_lo, int row_hi) {
    if (row_hi < 0) { row_lo = 0; row_hi = A->nrows; }
    if (row_hi <= row_lo || k <= 0) return CDG_OK;
    const StreamShape &s = A->shape64;
    if (s.T == 0) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "block size %d not supported", A->n);
    static const int cap = std::getenv("CDG_MV_MULTI_WIDTH") ? std::max(1, std::min(8, std::atoi(std::getenv("CDG_MV_MULTI_WIDTH")))) : 8;
    for (int r0 = 0; r0 < k; r0 += cap) {
        const int kk = std::min(cap, k - r0);
        const int K = kk <= 1 ? 1 : kk <= 2 ? 2 : kk <= 4 ? 4 : 8;
        const int chunk = std::max(1, std::min<int>(A->max_row_blocks, (32 * 1024) / (8 * A->n * K)));
        const size_t lds = sizeof(double) * ((size_t)chunk * K * A->n + 2 * (size_t)s.T);
        const double *x = X + (size_t)r0 * stride;
        double *y = Y + (size_t)r0 * stride;
        // loads in flight per thread: fewer as K grows (the products per load
        // grow with K; 16-byte loads x K accumulators must fit 128 VGPRs)
#define CDG_MVM(KK, UU) hipLaunchKernelGGL((k_mv_multi<KK, UU>), dim3(row_hi - row_lo), dim3(s.T), lds, ctx->stream, \
                                           A->blocks, A->ld, A->d_row_ptr, A->d_col, x, y, stride, kk, A->n, s.dc, chunk, \
                                           row_k, r0, row_lo)
        if (K == 1) CDG_MVM(1, 4);
        else if (K == 2) CDG_MVM(2, 4);
        else if (K == 4) CDG_MVM(4, 2);
        else CDG_MVM(8, 1);
#undef CDG_MVM
        CDG_HIP_TRY(hipGetLastError());
    }
    return CDG_OK;
}
