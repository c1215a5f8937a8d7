// ILU(0) factorization kernels for gfx950 (precon_ILU0.f90:111-265).
//
// Rows are processed in wavefront levels (rows of one level are independent).
// Per level:
//   phase 1  L_ij = A_ij * Dinv_j            batched GEMM on MFMA (one job per lower block)
//   phase 2  D_i  = A_ii - sum_j L_ij A_ji   batched GEMM on MFMA (one job per row)
//   phase 3  Dinv_i = inv(D_i)               one workgroup per block, register-resident
//                                            Gauss-Jordan with partial (row) pivoting
// FP64 uses v_mfma_f64_16x16x4f64, FP32 v_mfma_f32_16x16x4f32.  A 48x48 output
// tile per workgroup (3 waves, each 16x48 = three 16x16 MFMA tiles) covers the
// block sizes 40/135/320 with little padding.
#include "cdg_internal.h"

#include <algorithm>

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA.  Operands: lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Result row held by lane l in accumulator register r (column is l%16): the
// f64 instruction strides rows by 4 across registers, the f32 one keeps four
// consecutive rows per lane group (measured on gfx950).
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef f64x4 acc_t;
    __device__ static __forceinline__ acc_t run(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    typedef f32x4 acc_t;
    __device__ static __forceinline__ acc_t run(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int TM = 48;       // output tile edge
constexpr int KC = 16;       // K chunk staged in LDS
constexpr int GW = 3;        // waves per workgroup
constexpr int GT = 64 * GW;  // threads per workgroup

// acc += A[tm-tile rows, :] * B[:, tn-tile cols]; A, B column-major n x n.
template <typename T, typename SA, typename SB>
__device__ __forceinline__ void tile_mma(const SA *__restrict__ A, const SB *__restrict__ B, int n, int r0,
                                         int c0, T (*As)[TM], T (*Bs)[TM], typename Mfma<T>::acc_t *acc) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int kc = 0; kc < n; kc += KC) {
        for (int idx = tid; idx < KC * TM; idx += GT) {
            const int kk = idx / TM, rr = idx - kk * TM;
            const int gr = r0 + rr, gk = kc + kk;
            As[kk][rr] = (gr < n && gk < n) ? static_cast<T>(A[(size_t)gk * n + gr]) : T(0);
            const int cc = idx / KC, kb = idx - cc * KC;
            const int gc = c0 + cc, gk2 = kc + kb;
            Bs[kb][cc] = (gc < n && gk2 < n) ? static_cast<T>(B[(size_t)gc * n + gk2]) : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < KC / 4; ++s) {
            const T a = As[4 * s + (lane >> 4)][16 * wv + (lane & 15)];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const T b = Bs[4 * s + (lane >> 4)][16 * t + (lane & 15)];
                acc[t] = Mfma<T>::run(a, b, acc[t]);
            }
        }
        __syncthreads();
    }
}

// phase 1: L_q = A_{lo_blk[q]} * Dinv_{lo_col[q]}
template <typename T>
__global__ void __launch_bounds__(GT)
k_fac_p1(const double *__restrict__ Ab, int64_t lda, T *__restrict__ L, const T *__restrict__ D, int64_t ld,
         const int32_t *__restrict__ jobs, const int32_t *__restrict__ lo_blk, const int32_t *__restrict__ lo_col,
         int n, int tiles) {
    __shared__ T As[KC][TM], Bs[KC][TM];
    const int q = jobs[blockIdx.y];
    const int tm = blockIdx.x % tiles, tn = blockIdx.x / tiles;
    typename Mfma<T>::acc_t acc[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) acc[t] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    tile_mma<T, double, T>(Ab + (size_t)lo_blk[q] * lda, D + (size_t)lo_col[q] * ld, n, tm * TM, tn * TM, As, Bs, acc);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    T *out = L + (size_t)q * ld;
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = tm * TM + 16 * wv + Mfma<T>::row(lane, r);
            const int col = tn * TM + 16 * t + (lane & 15);
            if (row < n && col < n) out[(size_t)col * n + row] = acc[t][r];
        }
}

// phase 2: D_i = A_{diag[i]} - sum_q L_q * A_{lo_tr[q]}
template <typename T>
__global__ void __launch_bounds__(GT)
k_fac_p2(const double *__restrict__ Ab, int64_t lda, const T *__restrict__ L, T *__restrict__ D, int64_t ld,
         const int32_t *__restrict__ rows, const int32_t *__restrict__ lo_ptr, const int32_t *__restrict__ lo_tr,
         const int32_t *__restrict__ diag, int n, int tiles) {
    __shared__ T As[KC][TM], Bs[KC][TM];
    const int i = rows[blockIdx.y];
    const int tm = blockIdx.x % tiles, tn = blockIdx.x / tiles;
    typename Mfma<T>::acc_t acc[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) acc[t] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    for (int q = lo_ptr[i]; q < lo_ptr[i + 1]; ++q)
        tile_mma<T, T, double>(L + (size_t)q * ld, Ab + (size_t)lo_tr[q] * lda, n, tm * TM, tn * TM, As, Bs, acc);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const double *a = Ab + (size_t)diag[i] * lda;
    T *out = D + (size_t)i * ld;
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = tm * TM + 16 * wv + Mfma<T>::row(lane, r);
            const int col = tn * TM + 16 * t + (lane & 15);
            if (row < n && col < n) out[(size_t)col * n + row] = static_cast<T>(a[(size_t)col * n + row]) - acc[t][r];
        }
}

// phase 3: in-place inverse of D_i, one workgroup of 4 waves (one per SIMD)
// per block.  Thread (tr, tc) of the 16 x 16 thread grid keeps rows
// tr + 16a and columns tc + 16b (a, b < R) in registers.  Gauss-Jordan with
// partial pivoting and two barriers per pivot step:
//   1. the owners of column k publish it;
//   2. every wave finds the pivot row p itself (same data, same deterministic
//      reduction, so all waves agree; ties go to the lower row index) and the
//      owners of row p publish it;
//   3. everyone applies  S -= c_r * prow  (c_p = 0) and rewrites column k.
// Rows are never swapped; the final scatter inv(rowstep[r], pivrow[m]) undoes
// the implicit permutation.  The pivot row is not scaled in place: row r keeps
// its unscaled values and the factor sigma_r = 1/pivot is applied when the
// result is written (later eliminations are linear in the row, so they act on
// the unscaled values consistently).  Register indices are compile-time: the
// pivot loop runs k = 16*kb + kt with kb unrolled, and the pivot row is
// published through one LDS buffer per register row, so no register array is
// indexed at run time.
template <typename T, int R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_fac_inv(T *__restrict__ D, int64_t ld, const int32_t *__restrict__ rows, int n, int32_t *__restrict__ err) {
    constexpr int TR = 16;
    constexpr int NMAX = TR * R;
    constexpr int UR = (NMAX + 63) / 64;   // rows per lane in the pivot search
    __shared__ T colv[2][NMAX];
    __shared__ T rowv[2][R][NMAX];
    __shared__ T sig[NMAX];
    __shared__ int32_t pivrow[NMAX], rowstep[NMAX];
    const int blk = rows[blockIdx.x];
    T *M = D + (size_t)blk * ld;
    const int tid = threadIdx.x, tr = tid % TR, tc = tid / TR, lane = tid & 63;
    T S[R][R];
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b < R; ++b) {
            const int r = tr + TR * a, m = tc + TR * b;
            S[a][b] = (r < n && m < n) ? M[(size_t)m * n + r] : T(0);
        }
    for (int i = tid; i < 2 * NMAX; i += 256) (&colv[0][0])[i] = T(0);
    for (int i = tid; i < 2 * R * NMAX; i += 256) (&rowv[0][0][0])[i] = T(0);
    bool used[UR];
#pragma unroll
    for (int u = 0; u < UR; ++u) used[u] = false;
    int bad = 0;
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < R; ++kb) {
        const int kt_end = min(TR, n - kb * TR);
        for (int kt = 0; kt < kt_end; ++kt) {
            const int k = kb * TR + kt;
            const int buf = k & 1;
            if (tc == kt) {   // padding rows hold zeros, so no bounds test is needed
#pragma unroll
                for (int a = 0; a < R; ++a) colv[buf][tr + TR * a] = S[a][kb];
            }
            __syncthreads();
            // pivot: largest |a(r,k)| over unused rows, lowest index on ties
            T best = T(-1);
            int bi = 0x7fffffff, fu = 0x7fffffff;
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                const int r = lane + 64 * u;
                if (r < n && !used[u]) {
                    const T v = fabs(colv[buf][r]);
                    if (v > best) { best = v; bi = r; }
                    fu = min(fu, r);
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const T ob = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
                fu = min(fu, __shfl_xor(fu, off, 64));
            }
            int p = bi;
            if (!(best > T(0)) || p >= n) { bad = 1; p = fu; }   // zero or non-finite pivot column
            p = __builtin_amdgcn_readfirstlane(p);
#pragma unroll
            for (int u = 0; u < UR; ++u)
                if (lane + 64 * u == p) used[u] = true;
            const int pa = p / TR, pt = p - pa * TR;
            const T rinv = T(1) / colv[buf][p];
            if (tid == 0) { pivrow[k] = p; rowstep[p] = k; sig[p] = rinv; }
            if (tr == pt) {
#pragma unroll
                for (int a = 0; a < R; ++a)
                    if (a == pa) {
#pragma unroll
                        for (int b = 0; b < R; ++b) rowv[buf][a][tc + TR * b] = S[a][b];
                    }
            }
            __syncthreads();
            T prow[R], cr[R];
#pragma unroll
            for (int b = 0; b < R; ++b) prow[b] = rowv[buf][pa][tc + TR * b] * rinv;
#pragma unroll
            for (int a = 0; a < R; ++a) cr[a] = (tr + TR * a == p) ? T(0) : colv[buf][tr + TR * a];
#pragma unroll
            for (int a = 0; a < R; ++a)
#pragma unroll
                for (int b = 0; b < R; ++b) S[a][b] -= cr[a] * prow[b];
            // column k: -c_r/pivot, and 1 (times sigma_p = 1/pivot) in row p
#pragma unroll
            for (int a = 0; a < R; ++a) {
                const T v = (tr + TR * a == p) ? T(1) : -cr[a] * rinv;
                S[a][kb] = (tc == kt) ? v : S[a][kb];
            }
        }
    }
    __syncthreads();
    if (tid == 0 && bad) atomicMin(err, blk);
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b < R; ++b) {
            const int r = tr + TR * a, m = tc + TR * b;
            if (r < n && m < n) M[(size_t)pivrow[m] * n + rowstep[r]] = sig[r] * S[a][b];
        }
}

__device__ __forceinline__ cdg_bf16 to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return static_cast<cdg_bf16>(u >> 16);
}

template <typename S> __device__ __forceinline__ S narrow(double v);
template <> __device__ __forceinline__ float narrow<float>(double v) { return static_cast<float>(v); }
template <> __device__ __forceinline__ cdg_bf16 narrow<cdg_bf16>(double v) { return to_bf16(static_cast<float>(v)); }

// U_u = (S) A_{up_blk[u]}   (the apply's copy of A's upper blocks)
template <typename S>
__global__ void k_copy_upper(const double *__restrict__ Ab, int64_t lda, S *__restrict__ U, int64_t ld,
                             const int32_t *__restrict__ up_blk, int nup, int nn) {
    for (int u = blockIdx.y; u < nup; u += gridDim.y) {
        const double *src = Ab + (size_t)up_blk[u] * lda;
        S *dst = U + (size_t)u * ld;
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += gridDim.x * blockDim.x)
            dst[e] = narrow<S>(src[e]);
    }
}

__global__ void k_f32_to_bf16(const float *__restrict__ src, cdg_bf16 *__restrict__ dst, int64_t count) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x)
        dst[k] = to_bf16(src[k]);
}

template <typename T>
int launch_inv(cdg_ctx *ctx, T *D, int64_t ld, const int32_t *rows, int nrows_lvl, int n, int32_t *err) {
    dim3 grid(nrows_lvl), block(256);
    if (n <= 48) hipLaunchKernelGGL((k_fac_inv<T, 3>), grid, block, 0, ctx->stream, D, ld, rows, n, err);
    else if (n <= 96) hipLaunchKernelGGL((k_fac_inv<T, 6>), grid, block, 0, ctx->stream, D, ld, rows, n, err);
    else if (n <= 144) hipLaunchKernelGGL((k_fac_inv<T, 9>), grid, block, 0, ctx->stream, D, ld, rows, n, err);
    else if (n <= 192) hipLaunchKernelGGL((k_fac_inv<T, 12>), grid, block, 0, ctx->stream, D, ld, rows, n, err);
    else return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "block inverse for n=%d not built", n);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

template <typename T>
int factor_levels(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, T *L, T *D) {
    const int n = M->n;
    const int tiles = (n + TM - 1) / TM;
    for (int l = 0; l < M->nlev; ++l) {
        const int q0 = M->lev_lo_ptr[l], q1 = M->lev_lo_ptr[l + 1];
        const int r0 = M->lev_row_ptr[l], r1 = M->lev_row_ptr[l + 1];
        if (q1 > q0) {
            hipLaunchKernelGGL((k_fac_p1<T>), dim3(tiles * tiles, q1 - q0), dim3(GT), 0, ctx->stream, A->blocks, A->ld,
                               L, D, M->ld, M->d_lev_lo + q0, M->d_lo_blk, M->d_lo_col, n, tiles);
            CDG_HIP_TRY(hipGetLastError());
        }
        if (r1 > r0) {
            hipLaunchKernelGGL((k_fac_p2<T>), dim3(tiles * tiles, r1 - r0), dim3(GT), 0, ctx->stream, A->blocks, A->ld,
                               L, D, M->ld, M->d_lev_rows + r0, M->d_lo_ptr, M->d_lo_tr, M->d_diag, n, tiles);
            CDG_HIP_TRY(hipGetLastError());
            int e = launch_inv<T>(ctx, D, M->ld, M->d_lev_rows + r0, r1 - r0, n, M->d_err);
            if (e) return e;
        }
    }
    return CDG_OK;
}

}  // namespace

int launch_ilu_factor(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, int32_t *bad_row) {
    const int n = M->n;
    if (M->precision == CDG_FP64 && n > 192)
        return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "FP64 ILU factor for block size %d not built", n);
    if (n > 192)
        return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "ILU factor for block size %d not built", n);
    const int32_t big = 0x7fffffff;
    CDG_HIP_TRY(hipMemcpyAsync(M->d_err, &big, sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    int e = CDG_OK;
    if (M->precision == CDG_FP64) {
        e = factor_levels<double>(ctx, M, A, static_cast<double *>(M->L), static_cast<double *>(M->Dinv));
    } else if (M->precision == CDG_FP32) {
        e = factor_levels<float>(ctx, M, A, static_cast<float *>(M->L), static_cast<float *>(M->Dinv));
    } else {
        e = factor_levels<float>(ctx, M, A, M->L32, M->D32);
    }
    if (e) return e;
    const int nn = n * n;
    if (M->precision != CDG_FP64 && M->nup > 0) {
        dim3 grid((nn + 255) / 256 < 64 ? (nn + 255) / 256 : 64, std::min(M->nup, 65535));
        if (M->precision == CDG_FP32)
            hipLaunchKernelGGL(k_copy_upper<float>, grid, dim3(256), 0, ctx->stream, A->blocks, A->ld,
                               static_cast<float *>(M->U), M->ld, M->d_up_blk, M->nup, nn);
        else
            hipLaunchKernelGGL(k_copy_upper<cdg_bf16>, grid, dim3(256), 0, ctx->stream, A->blocks, A->ld,
                               static_cast<cdg_bf16 *>(M->U), M->ld, M->d_up_blk, M->nup, nn);
        CDG_HIP_TRY(hipGetLastError());
    }
    if (M->precision == CDG_BF16) {
        const int64_t cl = (int64_t)M->nlo * M->ld, cd = (int64_t)M->nrows * M->ld;
        if (cl) hipLaunchKernelGGL(k_f32_to_bf16, dim3(2048), dim3(256), 0, ctx->stream, M->L32,
                                   static_cast<cdg_bf16 *>(M->L), cl);
        if (cd) hipLaunchKernelGGL(k_f32_to_bf16, dim3(2048), dim3(256), 0, ctx->stream, M->D32,
                                   static_cast<cdg_bf16 *>(M->Dinv), cd);
        CDG_HIP_TRY(hipGetLastError());
    }
    int32_t err = big;
    CDG_HIP_TRY(hipMemcpyAsync(&err, M->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (err != big) {
        if (bad_row) *bad_row = err;
        return cdg_fail(ctx, CDG_ERR_SINGULAR, "singular diagonal block in row %d", err);
    }
    return CDG_OK;
}
