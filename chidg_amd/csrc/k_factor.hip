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

#include <type_traits>

#include <algorithm>
#include <cstdlib>
#include <mutex>

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

constexpr int KC = 16;       // K chunk staged in LDS

// Batched block GEMM.  A workgroup of TM / 16 waves computes a TM x TM tile
// (TM = 48 or 64) of one output block, acc = sum_p A_p B_p over the
// concatenated K of its operand pairs (A_p, B_p column-major n x n); wave w
// holds rows 16w..16w+15 (TM / 16 MFMA tiles of 16 x 16).  K advances in
// chunks of KC through two LDS buffers; the global fetch of chunk c + 2 goes
// out before chunk c's products, so a fetch has two chunks of products to
// arrive.  The 64-edge tile serves n > 192 (320 = 5 x 64: no padding, and 16
// products per wave between barriers instead of 12).
template <int TM> struct Geo {
    static constexpr int GW = TM / 16;   // waves per workgroup
    static constexpr int GT = 64 * GW;   // threads per workgroup
};
constexpr int LA = KC / 4;    // A elements per thread and chunk (KC * TM / GT)
constexpr int LB = KC / 4;    // B elements per thread and chunk
// LDS tiles are k-contiguous: As[r][k] = A(r0 + r, kc + k), Bs[c][k] =
// B(kc + k, c0 + c), pitch KP (16-byte rows; 20 floats spread the transposing
// A stores over the banks).  An MFMA sums over k, so lane l takes the
// contiguous k range (KC/4)(l/16) .. +KC/4 for both operands: one 16-byte LDS
// read per operand and step pair instead of one read per product.
template <typename T> struct KPitch { static constexpr int P = KC + 16 / (int)sizeof(T); };
template <typename T, int TM> using ATile = T[TM][KPitch<T>::P];
template <typename T, int TM> using BTile = T[TM][KPitch<T>::P];

// a 32-bit element offset from a uniform base (one VGPR per address)
template <typename S> __device__ __forceinline__ S ld_off(const S *base, int off) {
    return *reinterpret_cast<const S *>(reinterpret_cast<const char *>(base) + (uint32_t)(off * (int)sizeof(S)));
}

// The fetch loads raw source values (SA, SB) from in-bounds (clamped)
// addresses with no conditions, so nothing waits for them here; conversion to
// T and the zeros outside n happen when the chunk is stored to LDS.  (A
// conversion at fetch time makes every fetch wait for its own loads.)
template <typename SA, typename SB, int TM>
__device__ __forceinline__ void gemm_fetch(const SA *__restrict__ A, const SB *__restrict__ B, int n, int r0, int c0,
                                           int kc, SA (&ra)[LA], SB (&rb)[LB]) {
    constexpr int GT = Geo<TM>::GT;
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < LA; ++i) {
        const int idx = tid + GT * i;
        const int kk = idx / TM, rr = idx - kk * TM;
        ra[i] = ld_off(A, min(kc + kk, n - 1) * n + min(r0 + rr, n - 1));
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
        const int idx = tid + GT * i;
        const int cc = idx / KC, kb = idx - cc * KC;
        rb[i] = ld_off(B, min(c0 + cc, n - 1) * n + min(kc + kb, n - 1));
    }
}

template <typename T, typename SA, typename SB, int TM>
__device__ __forceinline__ void gemm_store(ATile<T, TM> &As, BTile<T, TM> &Bs, const SA (&ra)[LA], const SB (&rb)[LB],
                                           int n, int r0, int c0, int kc) {
    constexpr int GT = Geo<TM>::GT;
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < LA; ++i) {
        const int idx = tid + GT * i;
        const int kk = idx / TM, rr = idx - kk * TM;
        As[rr][kk] = (r0 + rr < n && kc + kk < n) ? static_cast<T>(ra[i]) : T(0);
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
        const int idx = tid + GT * i;
        const int cc = idx / KC, kb = idx - cc * KC;
        Bs[cc][kb] = (c0 + cc < n && kc + kb < n) ? static_cast<T>(rb[i]) : T(0);
    }
}

template <typename T, int TM>
__device__ __forceinline__ void gemm_chunk(const ATile<T, TM> &As, const BTile<T, TM> &Bs,
                                           typename Mfma<T>::acc_t (&acc)[TM / 16]) {
    constexpr int NJ = TM / 16;
    constexpr int KS = KC / 4;                   // k values per lane and chunk
    typedef T vec_t __attribute__((ext_vector_type(KS)));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k0 = KS * (lane >> 4);
    const vec_t a = *reinterpret_cast<const vec_t *>(&As[16 * wv + (lane & 15)][k0]);
    vec_t b[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) b[j] = *reinterpret_cast<const vec_t *>(&Bs[16 * j + (lane & 15)][k0]);
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[j] = Mfma<T>::run(a[s], b[j][s], acc[j]);
}

// acc = sum over the npairs operand pairs, pair(p, A, B) naming pair p.
template <typename T, typename SA, typename SB, int TM, typename Pair>
__device__ __forceinline__ void gemm_pairs(Pair pair, int npairs, int n, int r0, int c0, ATile<T, TM> (&As)[2],
                                           BTile<T, TM> (&Bs)[2], typename Mfma<T>::acc_t (&acc)[TM / 16]) {
    // Branch-free pipeline over an even number of chunks: fetches past the
    // last chunk reload it (in bounds) and their stores write zeros, so the
    // compiler's wait counts see loads in issue order (a conditional fetch
    // makes it wait for every outstanding load at each store).
    const int cpb = (n + KC - 1) / KC, nch = npairs * cpb, nch2 = (nch + 1) & ~1;
    SA ra0[LA], ra1[LA];
    SB rb0[LB], rb1[LB];
    auto fetch = [&](int c, SA (&ra)[LA], SB (&rb)[LB]) {
        c = min(c, nch - 1);
        const int p = c / cpb;
        const SA *A;
        const SB *B;
        pair(p, A, B);
        gemm_fetch<SA, SB, TM>(A, B, n, r0, c0, (c - p * cpb) * KC, ra, rb);
    };
    auto store = [&](int c, int buf, const SA (&ra)[LA], const SB (&rb)[LB]) {
        // chunk c >= nch: kc = n stores zeros
        const int kc = c < nch ? (c - (c / cpb) * cpb) * KC : n;
        gemm_store<T, SA, SB, TM>(As[buf], Bs[buf], ra, rb, n, r0, c0, kc);
    };
    fetch(0, ra0, rb0);
    fetch(1, ra1, rb1);
    store(0, 0, ra0, rb0);
    __syncthreads();
    for (int c = 0; c < nch2; c += 2) {
        fetch(c + 2, ra0, rb0);
        gemm_chunk<T, TM>(As[0], Bs[0], acc);
        store(c + 1, 1, ra1, rb1);
        __syncthreads();
        fetch(c + 3, ra1, rb1);
        gemm_chunk<T, TM>(As[1], Bs[1], acc);
        store(c + 2, 0, ra0, rb0);
        __syncthreads();
    }
}

// XCD-aware job/tile numbering.  Workgroups go to the 8 XCDs (each with its
// own L2) round-robin by linear id; id L runs tile (L / 8) % tiles2 of job
// 8 * ((L / 8) / tiles2) + L % 8, so all tiles of a job share one XCD's L2
// (they read the same operand strips).  Returns false for the ids past the
// last job.
__device__ __forceinline__ bool xcd_job(int njobs, int tiles2, int &job, int &tile) {
    const int L = blockIdx.x, k = L >> 3;
    job = 8 * (k / tiles2) + (L & 7);
    tile = k - (k / tiles2) * tiles2;
    return job < njobs;
}
inline int xcd_grid(int njobs, int tiles2) { return 8 * tiles2 * ((njobs + 7) / 8); }

template <typename T, int TM, typename F>
__device__ __forceinline__ void gemm_out(int tm, int tn, int n, typename Mfma<T>::acc_t (&acc)[TM / 16], F put) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < TM / 16; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = tm * TM + 16 * wv + Mfma<T>::row(lane, r);
            const int col = tn * TM + 16 * j + (lane & 15);
            if (row < n && col < n) put(row, col, acc[j][r]);
        }
}

// phase 1: L_q = A_{lo_blk[q]} * Dinv_{lo_col[q]} (with the upper lists in
// place of the lower ones: W_u = A_{up_blk[u]} * Dinv_{up_col[u]}); the A
// operand is A's own FP64 blocks or, for the FP32 fold, their FP32 copies
template <typename T, typename SA, int TM>
__global__ void __launch_bounds__(Geo<TM>::GT)
k_fac_p1(const SA *__restrict__ Ab, int64_t lda, T *__restrict__ L, const T *__restrict__ D, int64_t ld,
         const int32_t *__restrict__ jobs, int njobs, const int32_t *__restrict__ lo_blk,
         const int32_t *__restrict__ lo_col, int n, int tiles) {
    __shared__ __attribute__((aligned(16))) ATile<T, TM> As[2];
    __shared__ __attribute__((aligned(16))) BTile<T, TM> Bs[2];
    int job, tile;
    if (!xcd_job(njobs, tiles * tiles, job, tile)) return;
    const int q = jobs[job];
    const int tm = tile % tiles, tn = tile / tiles;
    typename Mfma<T>::acc_t acc[TM / 16];
#pragma unroll
    for (int j = 0; j < TM / 16; ++j) acc[j] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    const SA *A = Ab + (size_t)lo_blk[q] * lda;
    const T *B = D + (size_t)lo_col[q] * ld;
    gemm_pairs<T, SA, T, TM>([&](int, const SA *&a, const T *&b) { a = A; b = B; }, 1, n, tm * TM, tn * TM, As, Bs,
                             acc);
    T *out = L + (size_t)q * ld;
    gemm_out<T, TM>(tm, tn, n, acc, [&](int row, int col, T v) { out[(size_t)col * n + row] = v; });
}

// phase 2: D_i = A_{diag[i]} - sum_q L_q * U_q, U_q = A_{lo_tr[q]} read from
// Ub (stride ldu): A's own blocks (FP64), or their FP32 copies.
template <typename T, typename SB, int TM>
__global__ void __launch_bounds__(Geo<TM>::GT)
k_fac_p2(const double *__restrict__ Ab, int64_t lda, const SB *__restrict__ Ub, int64_t ldu, const T *__restrict__ L,
         T *__restrict__ D, int64_t ld, const int32_t *__restrict__ rows, int njobs, const int32_t *__restrict__ lo_ptr,
         const int32_t *__restrict__ lo_tr, const int32_t *__restrict__ diag, int n, int tiles) {
    __shared__ __attribute__((aligned(16))) ATile<T, TM> As[2];
    __shared__ __attribute__((aligned(16))) BTile<T, TM> Bs[2];
    int job, tile;
    if (!xcd_job(njobs, tiles * tiles, job, tile)) return;
    const int i = rows[job];
    const int tm = tile % tiles, tn = tile / tiles;
    typename Mfma<T>::acc_t acc[TM / 16];
#pragma unroll
    for (int j = 0; j < TM / 16; ++j) acc[j] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    const int q0 = lo_ptr[i], nq = lo_ptr[i + 1] - q0;
    if (nq > 0)
        gemm_pairs<T, T, SB, TM>(
            [&](int p, const T *&a, const SB *&b) {
                a = L + (size_t)(q0 + p) * ld;
                b = Ub + (size_t)lo_tr[q0 + p] * ldu;
            },
            nq, n, tm * TM, tn * TM, As, Bs, acc);
    const double *a = Ab + (size_t)diag[i] * lda;
    T *out = D + (size_t)i * ld;
    gemm_out<T, TM>(tm, tn, n, acc, [&](int row, int col, T v) {
        out[(size_t)col * n + row] = static_cast<T>(a[(size_t)col * n + row]) - v;
    });
}

// Storage conversion of a GEMM result (BF16: round to nearest even).
template <typename SO> __device__ __forceinline__ SO fold_store(double v);
template <> __device__ __forceinline__ double fold_store<double>(double v) { return v; }
template <> __device__ __forceinline__ float fold_store<float>(double v) { return static_cast<float>(v); }
template <> __device__ __forceinline__ cdg_bf16 fold_store<cdg_bf16>(double v) {
    uint32_t u = __float_as_uint(static_cast<float>(v));
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<cdg_bf16>(u >> 16);
}

// Row-slice fold: W_u = Dinv_{up_row[u]} * U_u for every upper entry u (the
// backward sweep's z_i = Dinv_i y_i - sum W_ik z_k), written straight into
// the row-slice layout of the stream schedule: row r of the block goes to
// slice q = r / w at element q*SL + col*w + (r - q*w) of the slot.  U_u is
// read from Ub slot up_src[u] (A's FP64 blocks, or their FP32 copies).
template <typename T, typename SB, typename SO, int TM>
__global__ void __launch_bounds__(Geo<TM>::GT)
k_fold_rs(const T *__restrict__ D, int64_t ld, const SB *__restrict__ Ub, int64_t ldu, const int32_t *__restrict__ up_row,
          const int32_t *__restrict__ up_src, int nup, SO *__restrict__ W, int64_t slot, int w, int SL, int n,
          int tiles) {
    __shared__ __attribute__((aligned(16))) ATile<T, TM> As[2];
    __shared__ __attribute__((aligned(16))) BTile<T, TM> Bs[2];
    int job, tile;
    if (!xcd_job(nup, tiles * tiles, job, tile)) return;
    const int u = job;
    const int tm = tile % tiles, tn = tile / tiles;
    typename Mfma<T>::acc_t acc[TM / 16];
#pragma unroll
    for (int j = 0; j < TM / 16; ++j) acc[j] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    const T *A = D + (size_t)up_row[u] * ld;
    const SB *B = Ub + (size_t)up_src[u] * ldu;
    gemm_pairs<T, T, SB, TM>([&](int, const T *&a, const SB *&b) { a = A; b = B; }, 1, n, tm * TM, tn * TM, As, Bs,
                             acc);
    SO *out = W + (size_t)u * slot;
    gemm_out<T, TM>(tm, tn, n, acc, [&](int row, int col, T v) {
        const int q = row / w;
        out[(size_t)q * SL + (size_t)col * w + (row - q * w)] = fold_store<SO>(static_cast<double>(v));
    });
}

// DPP moves inside a 16-lane row: xor 1, xor 2 (quad permutes), rotate 4, 8.
template <int CTRL> __device__ __forceinline__ int dpp_i32(int x) {
    return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ double dpp(double x) {
    const int2 v = __builtin_bit_cast(int2, x);
    return __builtin_bit_cast(double, int2{dpp_i32<CTRL>(v.x), dpp_i32<CTRL>(v.y)});
}
template <int CTRL> __device__ __forceinline__ float dpp(float x) {
    return __builtin_bit_cast(float, dpp_i32<CTRL>(__builtin_bit_cast(int, x)));
}

// (value, index) max with the lower index on ties, and a min of `fu`, over the
// 16 lanes of a DPP row; every lane of the row ends with the result.
template <typename T, int CTRL>
__device__ __forceinline__ void argmax_step(T &best, int &bi, int &fu) {
    const T ob = dpp<CTRL>(best);
    const int oi = dpp_i32<CTRL>(bi);
    fu = min(fu, dpp_i32<CTRL>(fu));
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
}

// phase 3: in-place inverse of D_i, one workgroup of 4 waves (one per SIMD)
// per block.  Thread (tr, tc) of the 16 x 16 thread grid keeps rows
// tr + 16a and columns tc + 16b (a, b < R) in registers; the 16 owners of a
// column are one 16-lane DPP row.  Gauss-Jordan with partial pivoting and two
// barriers per pivot step:
//   1. the owners of column k publish it and reduce it among themselves
//      (largest |a(r,k)| over unused rows, lower row index on ties);
//   2. everyone reads the pivot row p; its owners publish it;
//   3. everyone applies  S -= c_r * prow  (c_p = 0) and rewrites column k.
// Rows are never swapped; the final scatter inv(rowstep[r], pivrow[m]) undoes
// the implicit permutation.  The pivot row is not scaled in place: row r keeps
// its unscaled values and sigma_r = 1/pivot is applied when the result is
// written (later eliminations are linear in the row, so they act on the
// unscaled values consistently).  Register indices are compile-time: the
// pivot loop runs k = 16*kb + kt with kb unrolled, and the pivot row is
// published through one LDS buffer per register row.
template <typename T, int R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_fac_inv(T *__restrict__ D, int64_t ld, const int32_t *__restrict__ rows, int n, int32_t *__restrict__ err) {
    constexpr int TR = 16;
    constexpr int NMAX = TR * R;
    __shared__ T colv[2][NMAX];
    __shared__ T rowv[2][R][NMAX];
    __shared__ T sig[NMAX];
    __shared__ int32_t pivrow[NMAX], rowstep[NMAX];
    __shared__ int32_t piv_sh[2], bad_sh;
    const int blk = rows[blockIdx.x];
    T *M = D + (size_t)blk * ld;
    const int tid = threadIdx.x, tr = tid % TR, tc = tid / TR;
    T S[R][R];
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b < R; ++b) {
            const int r = tr + TR * a, m = tc + TR * b;
            S[a][b] = (r < n && m < n) ? M[(size_t)m * n + r] : T(0);
        }
    for (int i = tid; i < 2 * R * NMAX; i += 256) (&rowv[0][0][0])[i] = T(0);
    if (tid == 0) bad_sh = 0;
    bool used[R];
#pragma unroll
    for (int a = 0; a < R; ++a) used[a] = (tr + TR * a >= n);   // padding rows never pivot
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < R; ++kb) {
        const int kt_end = min(TR, n - kb * TR);
        for (int kt = 0; kt < kt_end; ++kt) {
            const int k = kb * TR + kt;
            const int buf = k & 1;
            if (tc == kt) {
                T best = T(-1);
                int bi = 0x7fffffff, fu = 0x7fffffff;
#pragma unroll
                for (int a = 0; a < R; ++a) {
                    const int r = tr + TR * a;
                    const T v = S[a][kb];
                    colv[buf][r] = v;   // padding rows hold zeros
                    if (!used[a]) {
                        const T av = fabs(v);
                        if (av > best) { best = av; bi = r; }
                        fu = min(fu, r);
                    }
                }
                argmax_step<T, 0xB1>(best, bi, fu);    // quad_perm [1,0,3,2]
                argmax_step<T, 0x4E>(best, bi, fu);    // quad_perm [2,3,0,1]
                argmax_step<T, 0x124>(best, bi, fu);   // row_ror:4
                argmax_step<T, 0x128>(best, bi, fu);   // row_ror:8
                if (tr == 0) {
                    if (!(best > T(0)) || bi >= n) { bad_sh = 1; bi = fu; }   // zero or non-finite column
                    piv_sh[buf] = bi;
                }
            }
            __syncthreads();
            const int p = __builtin_amdgcn_readfirstlane(piv_sh[buf]);
#pragma unroll
            for (int a = 0; a < R; ++a) used[a] = used[a] || (tr + TR * a == p);
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
    if (tid == 0 && bad_sh) atomicMin(err, blk);
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b < R; ++b) {
            const int r = tr + TR * a, m = tc + TR * b;
            if (r < n && m < n) M[(size_t)pivrow[m] * n + rowstep[r]] = sig[r] * S[a][b];
        }
}

// ---- blocked Gauss-Jordan inverse -------------------------------------------
// DPP with a row mask (lanes outside the mask keep their own value).
template <int CTRL, int RM> __device__ __forceinline__ int dpp_rm_i32(int x) {
    return __builtin_amdgcn_update_dpp(x, x, CTRL, RM, 0xF, false);
}
template <int CTRL, int RM> __device__ __forceinline__ double dpp_rm(double x) {
    const int2 v = __builtin_bit_cast(int2, x);
    return __builtin_bit_cast(double, int2{dpp_rm_i32<CTRL, RM>(v.x), dpp_rm_i32<CTRL, RM>(v.y)});
}
template <int CTRL, int RM> __device__ __forceinline__ float dpp_rm(float x) {
    return __builtin_bit_cast(float, dpp_rm_i32<CTRL, RM>(__builtin_bit_cast(int, x)));
}
template <typename T, int CTRL, int RM>
__device__ __forceinline__ void wargmax_step(T &best, int &bi, int &fu) {
    const T ob = dpp_rm<CTRL, RM>(best);
    const int oi = dpp_rm_i32<CTRL, RM>(bi);
    fu = min(fu, dpp_rm_i32<CTRL, RM>(fu));
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
}
// (value, index) maximum over the wave, lower index on ties, and the minimum
// of fu; the result is complete in lane 63.
template <typename T>
__device__ __forceinline__ void wave_argmax(T &best, int &bi, int &fu) {
    wargmax_step<T, 0xB1, 0xF>(best, bi, fu);    // quad_perm [1,0,3,2]
    wargmax_step<T, 0x4E, 0xF>(best, bi, fu);    // quad_perm [2,3,0,1]
    wargmax_step<T, 0x124, 0xF>(best, bi, fu);   // row_ror:4
    wargmax_step<T, 0x128, 0xF>(best, bi, fu);   // row_ror:8  (every lane: its row's result)
    wargmax_step<T, 0x142, 0xA>(best, bi, fu);   // row_bcast:15 into rows 1, 3
    wargmax_step<T, 0x143, 0xC>(best, bi, fu);   // row_bcast:31 into rows 2, 3
}
// Lanes of one wave exchanging data through LDS: without this the compiler
// may order another lane's load before the writing lane's store (each lane's
// program alone allows it).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ double readlane(double x, int l) {
    const int2 v = __builtin_bit_cast(int2, x);
    return __builtin_bit_cast(double, int2{__builtin_amdgcn_readlane(v.x, l), __builtin_amdgcn_readlane(v.y, l)});
}
__device__ __forceinline__ float readlane(float x, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}

// Diagonal-pivot threshold of the checked inverses: the diagonal a_kk is
// taken as the pivot of column k when |a_kk| >= kPivTau * |a_rk| for every
// unused row r (threshold partial pivoting with u = 0.1, the usual choice of
// sparse direct solvers: Duff, Erisman & Reid, Direct Methods for Sparse
// Matrices, ch. 5), which bounds the growth of a step by 1 + 1/kPivTau.  A
// block where some step fails the test is inverted with full partial
// pivoting instead.  Both give the inverse up to rounding; DG Jacobian
// blocks (dR/dq + mass/dtau) pass with |a_kk| / max_r |a_rk| >= 0.33 where
// strict partial pivoting would swap a third of the rows.
constexpr double kPivTau = 0.1;

// Gauss-Jordan with partial pivoting on one 16-column panel, run by one
// wave.  On entry G[r*gp + c] holds the panel (rows r < NP, zero beyond n and
// beyond the panel width w) and rowtag[r] != 0 marks rows pivoted in earlier
// panels.  On exit G holds C - E_P (C = [W on the pivot rows, -A(r,P) W
// elsewhere], W the inverse of the pivot block), rowtag[p] = 1 + step and
// pivrow[step] = p for this panel's pivots.  The pivot is the largest |a(r,k)|
// over unused rows, lower row on ties (DPP argmax); rows are never swapped.
// The panel's columns rotate left by one per step, so the column being
// eliminated always sits at register position 0 and the loop body is not
// unrolled; after 16 rotations the columns are back in place.
//
// PIVOT = false takes row k as the pivot of column k (the rows pivoted so far
// being 0..k-1) and only checks it against the threshold: no unused row with
// kPivTau * |a(r,k)| > |a(k,k)|, and a nonzero finite pivot.  The check is off
// the dependency chain of the elimination.  *miss reports a step where it did
// not hold.
template <typename T, int NR, bool PIVOT>
__device__ __forceinline__ bool panel_gj(T *G, int gp, int NP, int n, int k0, int w, int32_t *rowtag,
                                         int32_t *pivrow, T *prow_sh, bool *miss) {
    const int lane = threadIdx.x & 63;
    bool viol = false;
    T pr[NR][16];
    bool used[NR];
#pragma unroll
    for (int a = 0; a < NR; ++a) {
        const int r = lane + 64 * a;
        used[a] = r >= n || (r < NP && rowtag[r] != 0);
#pragma unroll
        for (int c = 0; c < 16; ++c) pr[a][c] = r < NP ? G[r * gp + c] : T(0);
    }
    bool bad = false;
#pragma unroll 1
    for (int j = 0; j < 16; ++j) {
        if (j < w) {
            int p = k0 + j;
            if (PIVOT) {
                T best = T(-1);
                int bi = 0x7fffffff, fu = 0x7fffffff;
#pragma unroll
                for (int a = 0; a < NR; ++a) {
                    const int r = lane + 64 * a;
                    if (!used[a]) {
                        const T av = fabs(pr[a][0]);
                        if (av > best) { best = av; bi = r; }
                        fu = min(fu, r);
                    }
                }
                wave_argmax<T>(best, bi, fu);
                const T bbest = readlane(best, 63);
                p = __builtin_amdgcn_readlane(bi, 63);
                if (!(bbest > T(0)) || p >= n) { bad = true; p = __builtin_amdgcn_readlane(fu, 63); }
            }
            const int pa = p >> 6, pl = p & 63;
            // the pivot row through LDS
            if (lane == pl) {
#pragma unroll
                for (int a = 0; a < NR; ++a)
                    if (a == pa) {
#pragma unroll
                        for (int c = 0; c < 16; ++c) prow_sh[c] = pr[a][c];
                    }
            }
            wave_sync();
            T prow[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) prow[c] = prow_sh[c];
            if (!PIVOT) {     // would partial pivoting have chosen row p?
                const T ap = fabs(prow[0]);
                if (!(ap > T(0)) || !(ap < T(INFINITY))) viol = true;
#pragma unroll
                for (int a = 0; a < NR; ++a)
                    if (!used[a] && (lane + 64 * a) != p && T(kPivTau) * fabs(pr[a][0]) > ap) viol = true;
            }
            const T rinv = T(1) / prow[0];
#pragma unroll
            for (int c = 0; c < 16; ++c) prow[c] *= rinv;
#pragma unroll
            for (int a = 0; a < NR; ++a) {
                const bool isp = (a == pa) && (lane == pl);
                const T cr = isp ? T(0) : pr[a][0];
#pragma unroll
                for (int c = 1; c < 16; ++c) pr[a][c] -= cr * prow[c];
                pr[a][0] = -cr * rinv;
                used[a] = used[a] || isp;
            }
            if (lane == pl) {
#pragma unroll
                for (int a = 0; a < NR; ++a)
                    if (a == pa) {
#pragma unroll
                        for (int c = 1; c < 16; ++c) pr[a][c] = prow[c];
                        pr[a][0] = rinv;
                    }
            }
            if (lane == 0) { rowtag[p] = k0 + j + 1; pivrow[k0 + j] = p; }
            wave_sync();   // prow_sh is rewritten by the next step; rowtag is read below
        }
#pragma unroll
        for (int a = 0; a < NR; ++a) {
            const T first = pr[a][0];
#pragma unroll
            for (int c = 0; c < 15; ++c) pr[a][c] = pr[a][c + 1];
            pr[a][15] = first;
        }
    }
    // C - E_P back to LDS (the ones of E_P sit at (p_j, j))
#pragma unroll
    for (int a = 0; a < NR; ++a) {
        const int r = lane + 64 * a;
        if (r < NP) {
            const int tg = rowtag[r] - 1 - k0;     // panel step of this row, if any
#pragma unroll
            for (int c = 0; c < 16; ++c) G[r * gp + c] = pr[a][c] - ((tg == c) ? T(1) : T(0));
        }
    }
    if (!PIVOT) *miss = __ballot(viol) != 0;
    return bad;
}

// The checked-diagonal panel of panel_gj<T, NR, false>, spread over the
// workgroup: row r sits in lane r of the block (NP <= 64 * NW), and waves past
// the last row only keep the barriers.  Step j: the owner of pivot row
// p = k0 + j checks its pivot, scales its row by 1 / a_pp (column j takes
// 1 / a_pp itself) and publishes it with |a_pp| through a double-buffered LDS
// row, one barrier per step (a wave can only overwrite buffer j % 2 after
// every wave has passed step j + 1's barrier, i.e. has read it); every other
// row r then takes a_rj times the scaled row (column j takes -a_rj / a_pp).
// The steps are unrolled, so each value keeps its register.  Outputs
// G = C - E_P, rowtag, pivrow; returns whether some step needed another pivot.
// Called by all 64 * NW threads.
template <typename T, int NW>
__device__ __forceinline__ bool panel_gjw(T *G, int gp, int NP, int n, int k0, int w, int32_t *rowtag,
                                          int32_t *pivrow, T (*prow_sh)[17], int32_t *miss_sh) {
    const int row = threadIdx.x;
    const bool live = row < NP;
    const bool wave_live = __builtin_amdgcn_readfirstlane(row & ~63) < NP;
    T pr[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) pr[c] = live ? G[row * gp + c] : T(0);
    bool used = row >= n || (live && rowtag[row] != 0);
    bool viol = false;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (j < w) {
            const int buf = j & 1;
            const bool isp = row == k0 + j;
            if (isp) {
                const T ap = fabs(pr[j]);
                if (!(ap > T(0)) || !(ap < T(INFINITY))) viol = true;
                const T rinv = T(1) / pr[j];
#pragma unroll
                for (int c = 0; c < 16; ++c) pr[c] = c == j ? rinv : pr[c] * rinv;
#pragma unroll
                for (int c = 0; c < 16; ++c) prow_sh[buf][c] = pr[c];
                prow_sh[buf][16] = ap;
            }
            __syncthreads();
            if (wave_live && !isp) {
                T prow[17];
#pragma unroll
                for (int c = 0; c < 17; ++c) prow[c] = prow_sh[buf][c];
                if (!used && T(kPivTau) * fabs(pr[j]) > prow[16]) viol = true;
                const T cr = pr[j];
#pragma unroll
                for (int c = 0; c < 16; ++c) pr[c] = c == j ? -cr * prow[j] : fma(-cr, prow[c], pr[c]);
            }
            used = used || isp;
        }
    }
    if (threadIdx.x < 16 && (int)threadIdx.x < w) { rowtag[k0 + threadIdx.x] = k0 + threadIdx.x + 1; pivrow[k0 + threadIdx.x] = k0 + threadIdx.x; }
    if (live) {
        const int tg = (row >= k0 && row < k0 + w) ? row - k0 : -1;   // this panel's pivot rows are k0..k0+w-1
#pragma unroll
        for (int c = 0; c < 16; ++c) G[row * gp + c] = pr[c] - ((tg == c) ? T(1) : T(0));
    }
    if (viol) *miss_sh = 1;
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*miss_sh) != 0;
}

// Lane gather (ds_bpermute: the value of lane addr / 4) and broadcast of quad
// lane Q to its quad (DPP quad_perm [Q,Q,Q,Q]).
__device__ __forceinline__ float bperm(float x, int addr) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double bperm(double x, int addr) {
    const int2 v = __builtin_bit_cast(int2, x);
    return __builtin_bit_cast(double, int2{__builtin_amdgcn_ds_bpermute(addr, v.x), __builtin_amdgcn_ds_bpermute(addr, v.y)});
}
template <int Q, typename T> __device__ __forceinline__ T quad_bcast(T x) { return dpp<Q * 0x55>(x); }

// 1/x for the FP32 pivot chain: v_rcp_f32 and one Newton step (within an ulp
// of the quotient; 3 dependent operations instead of the 11 of the IEEE
// division, which sits on the chain of every pivot step).  FP64 divides.
__device__ __forceinline__ float recip(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
}
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

// The checked-diagonal panel in two concurrent stages, for one workgroup of
// 64 * NW threads (NW >= 2).
//   Stage A (wave 0): the 16 x 16 pivot block (panel rows k0..k0+15), lane
//   4i + q holding row i, columns 4q..4q+3.  Step j takes the pivot a_jj with
//   v_readlane, the pivot row's entries with ds_bpermute and each row's
//   multiplier a_ij from its quad with DPP; the step's scaled pivot row s
//   (s_j = 1/a_jj) and the raw pivot go to the history H[j] (zeros for the
//   steps j >= w the panel does not have, which leave the rows unchanged), and
//   the step counter *step_sh is raised to base + j + 1 (release).
//   Stage B (waves 1..NW-1, one row per lane, row = lane * (NW-1) + wave - 1,
//   outside the pivot block) replays step j as soon as the counter shows it
//   (acquire), so the two stages overlap.  Wave 0 never waits, so every wait
//   ends.
// Both stages check the diagonal pivot of each step against the threshold:
// no unused row (row > k0 + j, row < n) with kPivTau * |a(r,j)| > |a(j,j)|,
// and a nonzero finite pivot.  The arithmetic is that of panel_gjw
// step for step.  On exit G holds C - E_P (E_P: ones at (k0 + c, c), c < w);
// returns whether some step needed another pivot.  base = 16 * (panel index):
// the counter only grows during the kernel.
constexpr int kHP = 20;   // history pitch (16 scaled pivot-row entries, raw pivot; 16-byte rows)
template <typename T, int NW>
__device__ __forceinline__ bool panel_ab(T *G, int gp, int NP, int n, int k0, int w, T (*H)[kHP], int32_t *miss_sh,
                                         int32_t *step_sh) {
    const int tid = threadIdx.x;
    const int base = k0;   // 16 steps per panel
    if (tid < 64) {
        const int i = tid >> 2, q = tid & 3;
        T a[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) a[e] = G[(k0 + i) * gp + 4 * q + e];
        bool viol = false;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int jq = j >> 2, je = j & 3;
            if (j < w) {
                const T piv = readlane(a[je], 4 * j + jq);
                const T rinv = recip(piv);
                const T ap = fabs(piv);
                if (!(ap > T(0)) || !(ap < T(INFINITY))) viol = true;
                T prow[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) prow[e] = bperm(a[e], (4 * j + q) * 4);
                T cr;
                switch (jq) {
                case 0: cr = quad_bcast<0>(a[je]); break;
                case 1: cr = quad_bcast<1>(a[je]); break;
                case 2: cr = quad_bcast<2>(a[je]); break;
                default: cr = quad_bcast<3>(a[je]); break;
                }
                if (i > j && k0 + i < n && T(kPivTau) * fabs(cr) > ap) viol = true;
                const bool isp = i == j, pc = q == jq;
                T sv[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) sv[e] = (pc && e == je) ? rinv : prow[e] * rinv;
                if (i == 0) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) H[j][4 * q + e] = sv[e];
                    if (q == 0) H[j][16] = piv;
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const T upd = isp ? sv[e] : fma(-cr, sv[e], a[e]);
                    a[e] = (e == je && pc) ? (isp ? rinv : -cr * rinv) : upd;
                }
            } else if (i == 0) {
#pragma unroll
                for (int e = 0; e < 4; ++e) H[j][4 * q + e] = T(0);
                if (q == 0) H[j][16] = T(0);
            }
            if (tid == 0) __hip_atomic_store(step_sh, base + j + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int c = 4 * q + e;
            G[(k0 + i) * gp + c] = a[e] - ((i == c && c < w) ? T(1) : T(0));
        }
        if (viol) *miss_sh = 1;
    } else {
        const int wv = tid >> 6, lane = tid & 63;
        const int row = lane * (NW - 1) + wv - 1;
        const bool live = row < NP && (row < k0 || row >= k0 + 16);
        const bool unused = row >= k0 + 16 && row < n;
        T pr[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) pr[c] = live ? G[row * gp + c] : T(0);
        bool viol = false;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            while (__hip_atomic_load(step_sh, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) <= base + j)
                __builtin_amdgcn_s_sleep(1);
            T sv[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) sv[c] = H[j][c];
            const T ap = fabs(H[j][16]);
            const T cr = pr[j];
            viol = viol | ((j < w) & unused & (T(kPivTau) * fabs(cr) > ap));
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (c != j) pr[c] = fma(-cr, sv[c], pr[c]);
            pr[j] = -cr * sv[j];
        }
        if (live) {
#pragma unroll
            for (int c = 0; c < 16; ++c) G[row * gp + c] = pr[c];
        }
        if (viol) *miss_sh = 1;
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*miss_sh) != 0;
}

// The checked-diagonal panel in two stages run one after the other, for one
// workgroup of 64 * NW threads (the FP64 form: panel_ab's concurrent stages
// need more registers than FP64 leaves).
//   Stage A (wave 0): the 16 x 16 pivot block (panel rows k0..k0+15), lane
//   4i + q holding row i, columns 4q..4q+3.  Step j takes the pivot a_jj with
//   v_readlane, the pivot row's entries with ds_bpermute and each row's
//   multiplier a_ij from its quad with DPP; the step's scaled pivot row s
//   (s_j = 1/a_jj) and the raw pivot go to the history H[j] (zeros for the
//   steps j >= w the panel does not have, which leave the rows unchanged).
//   Stage B (every lane, row lane * NW + wave outside the pivot block)
//   replays the 16 steps from H with no barrier between them.
// Both stages check the diagonal pivot of each step against the threshold:
// no unused row (row > k0 + j, row < n) with kPivTau * |a(r,j)| > |a(j,j)|,
// and a nonzero finite pivot.  The arithmetic is that of panel_gjw
// step for step.  On exit G holds C - E_P (E_P: ones at (k0 + c, c), c < w);
// returns whether some step needed another pivot.
template <typename T, int NW>
__device__ __forceinline__ bool panel_ab_seq(T *G, int gp, int NP, int n, int k0, int w, T (*H)[kHP], int32_t *miss_sh) {
    const int tid = threadIdx.x;
    if (tid < 64) {
        const int i = tid >> 2, q = tid & 3;
        T a[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) a[e] = G[(k0 + i) * gp + 4 * q + e];
        bool viol = false;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int jq = j >> 2, je = j & 3;
            if (j < w) {
                const T piv = readlane(a[je], 4 * j + jq);
                const T rinv = T(1) / piv;
                const T ap = fabs(piv);
                if (!(ap > T(0)) || !(ap < T(INFINITY))) viol = true;
                T prow[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) prow[e] = bperm(a[e], (4 * j + q) * 4);
                T cr;
                switch (jq) {
                case 0: cr = quad_bcast<0>(a[je]); break;
                case 1: cr = quad_bcast<1>(a[je]); break;
                case 2: cr = quad_bcast<2>(a[je]); break;
                default: cr = quad_bcast<3>(a[je]); break;
                }
                if (i > j && k0 + i < n && T(kPivTau) * fabs(cr) > ap) viol = true;
                const bool isp = i == j, pc = q == jq;
                T sv[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) sv[e] = (pc && e == je) ? rinv : prow[e] * rinv;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const T upd = isp ? sv[e] : fma(-cr, sv[e], a[e]);
                    a[e] = (e == je && pc) ? (isp ? rinv : -cr * rinv) : upd;
                }
                if (i == 0) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) H[j][4 * q + e] = sv[e];
                    if (q == 0) H[j][16] = piv;
                }
            } else if (i == 0) {
#pragma unroll
                for (int e = 0; e < 4; ++e) H[j][4 * q + e] = T(0);
                if (q == 0) H[j][16] = T(0);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int c = 4 * q + e;
            G[(k0 + i) * gp + c] = a[e] - ((i == c && c < w) ? T(1) : T(0));
        }
        if (viol) *miss_sh = 1;
    }
    __syncthreads();
    const int wv = tid >> 6, lane = tid & 63;
    const int row = lane * NW + wv;
    const bool live = row < NP && (row < k0 || row >= k0 + 16);
    const bool unused = row >= k0 + 16 && row < n;
    T pr[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) pr[c] = live ? G[row * gp + c] : T(0);
    bool viol = false;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        T sv[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) sv[c] = H[j][c];
        const T ap = fabs(H[j][16]);
        const T cr = pr[j];
        viol = viol | ((j < w) & unused & (T(kPivTau) * fabs(cr) > ap));
#pragma unroll
        for (int c = 0; c < 16; ++c)
            if (c != j) pr[c] = fma(-cr, sv[c], pr[c]);
        pr[j] = -cr * sv[j];
    }
    if (live) {
#pragma unroll
        for (int c = 0; c < 16; ++c) G[row * gp + c] = pr[c];
    }
    if (viol) *miss_sh = 1;
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*miss_sh) != 0;
}

// Raw buffer access to one block's slot: a byte offset at or beyond the
// descriptor's size (kOob) reads 0 and drops the store, so masked lanes need
// no branch (descriptor word 3: raw, untyped access on CDNA).
typedef unsigned int bu32x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kOob = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slot_rsrc(void *p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, bytes, 0x00020000);
}
template <typename T> __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <typename T> __device__ __forceinline__ void bstore(T v, __amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ __forceinline__ void bstore<float>(float v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, off, 0, 0);
}
template <> __device__ __forceinline__ void bstore<double>(double v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bu32x2, v), r, off, 0, 0);
}

// Blocked Gauss-Jordan inverse with partial pivoting of the n x n block M
// (n <= 64*NR) by one workgroup, working in the global slot S (L2-resident):
// the rare blocks that the diagonal-pivot inverse hands over.  The rank-16
// update runs on the transposed tiles, D'(c, r) = S(r, c): an MFMA tile row
// is then a matrix column, so its 16 lanes touch 16 consecutive elements of
// the column-major block.  LDS: the panel G (NP x 17), the pivot rows XP
// (16 x (NP + 1)), prow_sh, rowtag, pivrow (NP = 16 * ceil(n/16)).  M
// receives the inverse and *bad reports a zero or non-finite pivot column.
template <typename T, int NR, int NTH>
__device__ void ginv_pivot_body(T *__restrict__ M, T *__restrict__ S, int n, T *G, T *XP, T *prow_sh,
                                int32_t *rowtag, int32_t *pivrow, int32_t *flag_sh, bool *bad) {
    constexpr int GP = 17, NWV = NTH / 64, UBP = 4;
    typedef typename Mfma<T>::acc_t acc_t;
    const int NT = (n + 15) / 16, NP = 16 * NT, XPP = NP + 1;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nn = n * n;
    for (int e = tid; e < nn; e += NTH) S[e] = M[e];
    for (int i = tid; i < NP; i += NTH) { rowtag[i] = 0; pivrow[i] = 0; }
    for (int i = tid; i < 16 * XPP; i += NTH) XP[i] = T(0);
    if (tid == 0) *flag_sh = 0;
    __syncthreads();
    for (int kb = 0; kb < NT; ++kb) {
        const int k0 = 16 * kb;
        const int w = min(16, n - k0);
        // 1. panel to LDS (zero beyond n and beyond the panel width)
        for (int e = tid; e < NP * 16; e += NTH) {
            const int c = e / NP, r = e - c * NP;
            G[r * GP + c] = (r < n && c < w) ? S[(size_t)(k0 + c) * n + r] : T(0);
        }
        __syncthreads();
        // 2. panel Gauss-Jordan with partial pivoting (wave 0); flag: zero pivot
        if (wv == 0) {
            bool miss = false;
            const bool b = panel_gj<T, NR, true>(G, GP, NP, n, k0, w, rowtag, pivrow, prow_sh, &miss);
            if (lane == 0 && b) *flag_sh = 1;
        }
        __syncthreads();
        // 3. pivot rows X(P, :) of the other columns
        for (int e = tid; e < w * n; e += NTH) {
            const int c = e / w, j = e - c * w;
            if (c < k0 || c >= k0 + w) XP[j * XPP + c] = S[(size_t)c * n + pivrow[k0 + j]];
        }
        __syncthreads();
        // 4. X += (C - E_P) X(P,:) on MFMA (transposed tiles) for the tiles outside
        //    the panel's tile column, then the panel's tiles take C
        for (int t0 = wv; t0 < NT * NT; t0 += NWV * UBP) {
            acc_t acc[UBP];
#pragma unroll
            for (int u = 0; u < UBP; ++u) {
                const int t = t0 + NWV * u;
                const int ti = t % NT, tj = t / NT;
                const int r = 16 * ti + (lane & 15);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int c = 16 * tj + Mfma<T>::row(lane, q);
                    acc[u][q] = (t < NT * NT && tj != kb && r < n && c < n) ? S[(size_t)c * n + r] : T(0);
                }
            }
#pragma unroll
            for (int u = 0; u < UBP; ++u) {
                const int t = t0 + NWV * u;
                if (t >= NT * NT) continue;
                const int ti = t % NT, tj = t / NT;
                const int r = 16 * ti + (lane & 15);
                if (tj != kb) {
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        const T a = XP[(4 * kk + (lane >> 4)) * XPP + 16 * tj + (lane & 15)];
                        const T bb = G[(16 * ti + (lane & 15)) * GP + 4 * kk + (lane >> 4)];
                        acc[u] = Mfma<T>::run(a, bb, acc[u]);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int c = 16 * tj + Mfma<T>::row(lane, q);
                        if (r < n && c < n) S[(size_t)c * n + r] = acc[u][q];
                    }
                } else if (r < n) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int cc = Mfma<T>::row(lane, q);
                        const int c = k0 + cc;
                        if (cc < w) S[(size_t)c * n + r] = G[r * GP + cc] + ((rowtag[r] - 1 - k0 == cc) ? T(1) : T(0));
                    }
                }
            }
        }
        __syncthreads();
    }
    *bad = __builtin_amdgcn_readfirstlane(*flag_sh) != 0;
    // inv(rowstep(r), pivrow(c)) = S(r, c)
    for (int e = tid; e < nn; e += NTH) {
        const int c = e / n, r = e - c * n;
        M[(size_t)pivrow[c] * n + (rowtag[r] - 1)] = S[e];
    }
}

// Wide-panel inverse with the pivot rows in registers: panels of H 16-column
// halves in LDS (each half's Gauss-Jordan steps, panel_gjw, then their
// transform on the other halves on MFMA), then every tile outside the panel
// takes the rank-16H update.
// Each wave owns whole tile columns; before updating any tile of a column it
// loads that column's pivot-row operands (16H rows, 4H values per lane) into
// registers, so no other wave can have changed them (the column's tiles, its
// pivot-row tiles among them, are only this wave's).  Fewer, wider panels
// mean fewer passes over the slot (each reads and writes the whole block);
// with the pivot rows out of LDS, the halves alone take it.  The first panel
// reads M and the last writes M (the slot S holds the passes between), so
// neither a copy of M nor a copy back is needed; the result is inv(M) with
// no permutation, as every pivot is diagonal.  LDS (ginv_lds): H halves of
// NP x 17, rowtag, pivrow.
template <typename T> struct GinvHalves { static constexpr int H = sizeof(T) == 4 ? 4 : 3; };
template <typename T, int NR, int NTH, int UB, int H>
__device__ bool ginv_body_wide_rx(T *__restrict__ M, T *__restrict__ S, int n, T *G, T (*prow2_sh)[17],
                                  int32_t *rowtag, int32_t *pivrow, int32_t *flag_sh) {
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int GP = 17, NWV = NTH / 64, PW = 16 * H;
    const int NT = (n + 15) / 16, NP = 16 * NT;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nn = n * n;
    auto half = [&](int h) -> T * { return G + (size_t)h * NP * GP; };
    for (int i = tid; i < NP; i += NTH) { rowtag[i] = 0; pivrow[i] = 0; }
    if (tid == 0) *flag_sh = 0;
    __syncthreads();
    // the first panel reads the block, the last one writes it, the others work in the
    // slot (M stays as it was until the last panel's pivots have passed their check)
    const __amdgpu_buffer_rsrc_t rsM = slot_rsrc(M, nn * (int)sizeof(T)), rsS = slot_rsrc(S, nn * (int)sizeof(T));
    constexpr int PL0 = (PW * 64 * NR + NTH - 1) / NTH;   // panel values per thread
    constexpr int PB = PL0 < 10 ? PL0 : 10;   // loaded per batch (registers)
    constexpr int PL = (PL0 + PB - 1) / PB * PB;
    for (int kb = 0; kb < NT; kb += H) {
        const int k0 = 16 * kb;
        const int np = min(H, NT - kb);                         // halves (tile columns) of this panel
        auto width = [&](int h) { return h < np ? min(16, n - k0 - 16 * h) : 0; };
        const __amdgpu_buffer_rsrc_t src = kb == 0 ? rsM : rsS, dst = kb + H >= NT ? rsM : rsS;
        // 1. the halves to LDS (zero beyond n and beyond each half's width)
#pragma unroll 1
        for (int i0 = 0; i0 < PL; i0 += PB) {
            T v[PB];
#pragma unroll
            for (int i = 0; i < PB; ++i) {
                const int e = tid + (i0 + i) * NTH, c = e / NP, r = e - c * NP;
                const bool ok = e < NP * PW && r < n && (c & 15) < width(c >> 4);
                v[i] = bload<T>(src, ok ? (uint32_t)(((k0 + c) * n + r) * (int)sizeof(T)) : kOob);
            }
#pragma unroll
            for (int i = 0; i < PB; ++i) {
                const int e = tid + (i0 + i) * NTH, c = e / NP, r = e - c * NP;
                if (e < NP * PW) half(c >> 4)[r * GP + (c & 15)] = v[i];
            }
        }
        __syncthreads();
        // 2. each half's steps, then their transform on the panel's other halves t:
        //    G_t += G_h G_t(P_h, :) on MFMA tiles (P_h: half h's pivot rows), whose
        //    pivot-row operands every wave reads before any tile changes
        for (int h = 0; h < np; ++h) {
            const int kh = k0 + 16 * h, wh = width(h);
            if (panel_gjw<T, NWV>(half(h), GP, NP, n, kh, wh, rowtag, pivrow, prow2_sh, flag_sh)) return false;
            if (np == 1) break;
            T b[H][4];
#pragma unroll
            for (int t = 0; t < H; ++t)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int j = 4 * kk + (lane >> 4);
                    b[t][kk] = (t < np && t != h && j < wh) ? half(t)[pivrow[kh + j] * GP + (lane & 15)] : T(0);
                }
            __syncthreads();
            const T *Gs = half(h);
            for (int idx = wv; idx < NT * (np - 1); idx += NWV) {
                const int s = idx / NT, ti = idx - s * NT, t = s + (s >= h);
                T bs[4];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    bs[kk] = b[0][kk];
#pragma unroll
                    for (int u = 1; u < H; ++u) bs[kk] = t == u ? b[u][kk] : bs[kk];
                }
                T *Gt = half(t);
                acc_t acc;
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = Gt[(16 * ti + Mfma<T>::row(lane, q)) * GP + (lane & 15)];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
                    acc = Mfma<T>::run(Gs[(16 * ti + (lane & 15)) * GP + 4 * kk + (lane >> 4)], bs[kk], acc);
#pragma unroll
                for (int q = 0; q < 4; ++q) Gt[(16 * ti + Mfma<T>::row(lane, q)) * GP + (lane & 15)] = acc[q];
            }
            __syncthreads();
        }
        // 3. X += (C - E_P) X(P, :) on the tiles outside the panel (transposed tiles,
        //    branch-free, the next round's loads before this round's stores), then the
        //    panel's tiles take C.  A wave's work item is a tile column, or half of one
        //    when the columns do not divide evenly among the waves (17 columns on 16
        //    waves: three rounds of halves instead of two of columns).  Both waves of a
        //    split column read its pivot rows before either changes them (one barrier per
        //    round; a column's two halves are items 2j, 2j + 1 of one round).
        const int ncol = NT - np;
        const int split = ncol % NWV != 0 ? 2 : 1;
        const int nh = split == 2 ? (NT + 1) / 2 : NT;     // tiles in a column's first half
        for (int it0 = 0; it0 < split * ncol; it0 += NWV) {
            const int item = it0 + wv;
            const bool act = item < split * ncol;           // uniform per wave
            const int tjp = min(item, split * ncol - 1) / split, part = item - tjp * split;
            const int tj = tjp + (tjp >= kb ? np : 0);
            const int lo = part == 1 ? nh : 0, hi = part == 1 || split == 1 ? NT : nh;
            const int c = 16 * tj + (lane & 15);
            T a[4 * H];   // X(P, c): pivot row 4 kk + lane / 16 of the panel, this lane's column
#pragma unroll
            for (int kk = 0; kk < 4 * H; ++kk) {
                const int h = kk >> 2, j = 4 * (kk & 3) + (lane >> 4);
                const bool ok = act && h < np && j < width(h) && c < n;
                a[kk] = bload<T>(src, ok ? (uint32_t)((c * n + pivrow[min(k0 + 16 * h + j, NP - 1)]) * (int)sizeof(T)) : kOob);
            }
            if (split == 2) __syncthreads();
            if (!act) continue;
            auto offs = [&](int ti, int q) -> uint32_t {
                const int r = 16 * ti + (lane & 15), cq = 16 * tj + Mfma<T>::row(lane, q);
                return (ti < hi && r < n && cq < n) ? (uint32_t)((cq * n + r) * (int)sizeof(T)) : kOob;
            };
            auto load_round = [&](T (&v)[UB][4], int i0) {
#pragma unroll
                for (int u = 0; u < UB; ++u)
#pragma unroll
                    for (int q = 0; q < 4; ++q) v[u][q] = bload<T>(src, offs(i0 + u, q));
            };
            auto update_round = [&](const T (&v)[UB][4], int i0) {
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    const int ti = min(i0 + u, NT - 1);
                    const T *grow = G + (16 * ti + (lane & 15)) * GP + (lane >> 4);
                    acc_t acc;
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] = v[u][q];
#pragma unroll
                    for (int kk = 0; kk < 4 * H; ++kk) {
                        if (kk < 4 * np) {
                            const T bb = grow[(kk >> 2) * NP * GP + 4 * (kk & 3)];
                            acc = Mfma<T>::run(a[kk], bb, acc);
                        }
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) bstore<T>(acc[q], dst, offs(i0 + u, q));
                }
            };
            T va[UB][4], vb[UB][4];
            load_round(va, lo);
            for (int i0 = lo; i0 < hi; i0 += 2 * UB) {
                load_round(vb, i0 + UB);
                update_round(va, i0);
                load_round(va, i0 + 2 * UB);
                if (i0 + UB < hi) update_round(vb, i0 + UB);
            }
        }
        for (int h = 0; h < np; ++h) {
            const T *Gh = half(h);
            const int kh = k0 + 16 * h, wh = width(h);
            for (int ti = wv; ti < NT; ti += NWV) {
                const int r = 16 * ti + (lane & 15);
                if (r < n) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int cc = Mfma<T>::row(lane, q);
                        if (cc < wh)
                            bstore<T>(Gh[r * GP + cc] + ((rowtag[r] - 1 - kh == cc) ? T(1) : T(0)), dst,
                                      (uint32_t)(((kh + cc) * n + r) * (int)sizeof(T)));
                    }
                }
            }
        }
        __syncthreads();
    }
    return true;   // every pivot was diagonal: the last panel left inv(M) in place
}

// Dinv_i = inv(D_i) for any n <= 64*NR: ginv_body_wide_rx with diagonal
// pivots; a block whose check fails is flagged for k_fac_pivot.  LDS is dynamic.
constexpr int kInvThreads = 512;    // partial-pivoting inverse (k_fac_pivot): 8 waves per block
constexpr int kWideThreads = 1024;   // wide-panel inverse: 16 waves per block

template <typename T, int NR, int UB>
__global__ void __launch_bounds__(kWideThreads) k_fac_ginv(T *__restrict__ D, int64_t ld, T *__restrict__ Wk, int64_t ldw,
                                                   const int32_t *__restrict__ rows, int n,
                                                   int32_t *__restrict__ redo) {
    extern __shared__ unsigned char smem_raw[];
    const int NP = 16 * ((n + 15) / 16);
    __shared__ int32_t flag_sh;
    __shared__ T prow2_sh[2][17];
    const int blk = rows[blockIdx.x];
    T *M = D + (size_t)blk * ld;
    T *S = Wk + (size_t)blockIdx.x * ldw;
    // panels of GinvHalves<T>::H 16-column halves, pivot rows in registers (LDS: ginv_lds)
    constexpr int H = GinvHalves<T>::H;
    T *G = reinterpret_cast<T *>(smem_raw);                 // H * NP * 17
    int32_t *rowtag = reinterpret_cast<int32_t *>(G + (size_t)H * NP * 17);   // NP
    int32_t *pivrow = rowtag + NP;                          // NP
    const bool ok = ginv_body_wide_rx<T, NR, kWideThreads, UB, H>(M, S, n, G, prow2_sh, rowtag, pivrow, &flag_sh);
    if (!ok && threadIdx.x == 0) redo[blk] = 1;    // left to k_fac_pivot; M untouched
}

// Dinv_i = inv(D_i) for blocks of n <= 16*NT, one workgroup of 4 waves per
// block, pivoting on the diagonal.  The block lives in MFMA accumulator tiles.
// Wave v holds the whole tile columns tj = v + 4c (c < NT / 4), so that those
// tiles' operands sit at compile-time offsets from a per-lane base; the tiles
// of the last NT % 4 columns are dealt round-robin (tile e to wave e % 4), so
// that no wave holds more than ceil(NT^2 / 4) tiles.  Columns are processed in
// panels of 16:
//   1. the panel's tiles go to LDS (G), and so do the other tiles of the
//      panel's pivot rows k0..k0+15 (XP);
//   2. panel_ab (FP32) or panel_ab_seq (FP64) runs the checked-diagonal
//      Gauss-Jordan on the n x 16 panel,
//      which then holds C = [W on the pivot rows P, -A(r,P) W elsewhere], W
//      the inverse of the panel's pivot block;
//   3. every other column is updated at once, X += (C - E_P) X(P,:), one
//      rank-16 product on MFMA (E_P: ones at (k0 + j, j)), and the panel's
//      tiles take C, formed on MFMA as (C - E_P + E_P) I from a zero
//      accumulator (an accumulator written by VALU makes the compiler hold
//      every tile in arch VGPRs and copy them all in each panel).
// This is the unblocked in-place Gauss-Jordan reorganized, on diagonal pivots
// that pass the kPivTau threshold; a block where one fails is left untouched
// and flagged in redo[] for k_fac_pivot.
template <typename T, int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_fac_binv(T *__restrict__ D, int64_t ld, const int32_t *__restrict__ rows, int n, int32_t *__restrict__ err,
           int32_t *__restrict__ redo) {
    constexpr int NP = 16 * NT;                  // padded size
    constexpr int NFC = NT / 4;                  // whole tile columns per wave
    constexpr int NFA = NFC > 0 ? NFC : 1;
    constexpr int NXT = (NT - 4 * NFC) * NT;     // tiles of the remaining columns
    constexpr int NX = (NXT + 3) / 4;            // of which one wave holds at most NX
    constexpr int NXA = NX > 0 ? NX : 1;
    constexpr int GP = 17;                       // LDS pitch of the panel (bank spread)
    constexpr int XPP = NP + 1;                  // LDS pitch of the pivot rows
    __shared__ T G[NP * GP];                     // panel in, C - E_P out
    __shared__ T XP[16 * XPP];                   // pivot rows of the other columns
    __shared__ T H[16][kHP];                     // panel step history
    __shared__ int32_t miss_sh, step_sh;
    typedef typename Mfma<T>::acc_t acc_t;
    const int blk = rows[blockIdx.x];
    T *M = D + (size_t)blk * ld;
    const int tid = threadIdx.x, lane0 = tid & 63;
    const int wv0 = __builtin_amdgcn_readfirstlane(tid >> 6);
    acc_t acc[NFA][NT];     // tile (ti, wave + 4c)
    acc_t accx[NXA];        // remaining tile e = wave + 4x: (e % NT, 4 NFC + e / NT)
    // a 32-bit byte offset from the block's (uniform) base keeps one VGPR per address
    auto ldM = [&](int row, int col, bool ok) -> T {
        return (ok && row < n && col < n) ? *reinterpret_cast<const T *>(
                   reinterpret_cast<const char *>(M) + (uint32_t)((col * n + row) * (int)sizeof(T))) : T(0);
    };
#pragma unroll
    for (int c = 0; c < NFC; ++c)
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                acc[c][ti][r] = ldM(16 * ti + Mfma<T>::row(lane0, r), 16 * (wv0 + 4 * c) + (lane0 & 15), true);
#pragma unroll
    for (int x = 0; x < NX; ++x) {
        const int e = wv0 + 4 * x, eti = e % NT, etj = 4 * NFC + e / NT;
#pragma unroll
        for (int r = 0; r < 4; ++r) accx[x][r] = ldM(16 * eti + Mfma<T>::row(lane0, r), 16 * etj + (lane0 & 15), e < NXT);
    }
    for (int e = tid; e < 16 * XPP; e += 256) XP[e] = T(0);
    if (tid == 0) { miss_sh = 0; step_sh = 0; }
    // B operand of the identity: I[4kk + lane/16][lane%16]
    T ident[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) ident[kk] = (lane0 & 15) == 4 * kk + (lane0 >> 4) ? T(1) : T(0);
    __syncthreads();
    for (int kb = 0; kb < NT; ++kb) {
        const int k0 = 16 * kb;
        const int w = min(16, n - k0);
        // wave and lane indices made opaque per panel: the LDS addresses below
        // are recomputed in each panel rather than hoisted out of the loop
        // (kept live across it they exceed the register file)
        int wv, lane;
        asm volatile("s_mov_b32 %0, %1" : "=s"(wv) : "s"(wv0));
        asm volatile("v_mov_b32 %0, %1" : "=v"(lane) : "v"(lane0));
        auto storeG = [&](const acc_t &v, int ti) {
#pragma unroll
            for (int r = 0; r < 4; ++r) G[(16 * ti + Mfma<T>::row(lane, r)) * GP + (lane & 15)] = v[r];
        };
        auto storeXP = [&](const acc_t &v, int tj) {
#pragma unroll
            for (int r = 0; r < 4; ++r) XP[Mfma<T>::row(lane, r) * XPP + 16 * tj + (lane & 15)] = v[r];
        };
        // 1. the panel's tile column to G, the pivot rows' other tiles to XP
#pragma unroll
        for (int c = 0; c < NFC; ++c) {
            const int tj = wv + 4 * c;
            if (tj == kb) {
#pragma unroll
                for (int ti = 0; ti < NT; ++ti) storeG(acc[c][ti], ti);
            } else {
#pragma unroll
                for (int ti = 0; ti < NT; ++ti)
                    if (ti == kb) storeXP(acc[c][ti], tj);
            }
        }
#pragma unroll
        for (int x = 0; x < NX; ++x) {
            const int e = wv + 4 * x, eti = e % NT, etj = 4 * NFC + e / NT;
            if (e < NXT) {
                if (etj == kb) storeG(accx[x], eti);
                else if (eti == kb) storeXP(accx[x], etj);
            }
        }
        __syncthreads();
        // the pivot rows (B operands of step 3) are final now
        T b[NFA][4], bx[NXA][4];
#pragma unroll
        for (int c = 0; c < NFC; ++c)
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) b[c][kk] = XP[(4 * kk + (lane >> 4)) * XPP + 16 * (wv + 4 * c) + (lane & 15)];
#pragma unroll
        for (int x = 0; x < NX; ++x) {
            const int e = min(wv + 4 * x, NXT - 1), etj = 4 * NFC + e / NT;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) bx[x][kk] = XP[(4 * kk + (lane >> 4)) * XPP + 16 * etj + (lane & 15)];
        }
        // 2. panel Gauss-Jordan (checked diagonal pivots)
        const bool miss = sizeof(T) == 4 ? panel_ab<T, 4>(G, GP, NP, n, k0, w, H, &miss_sh, &step_sh)
                                         : panel_ab_seq<T, 4>(G, GP, NP, n, k0, w, H, &miss_sh);
        if (miss) break;
        // 3. X += (C - E_P) X(P,:) on MFMA; the panel's tiles take C
        auto ldA = [&](int ti, int kk) -> T { return G[(16 * ti + (lane & 15)) * GP + 4 * kk + (lane >> 4)]; };
        auto panel_tile = [&](const T (&av)[4], int ti) -> acc_t {
            acc_t r = {};
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int cc = 4 * kk + (lane >> 4);
                const T ae = av[kk] + ((ti == kb && (lane & 15) == cc && cc < w) ? T(1) : T(0));
                r = Mfma<T>::run(ae, ident[kk], r);
            }
            return r;
        };
        if (NFC > 0) {
            T a[NT][4];
#pragma unroll
            for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) a[ti][kk] = ldA(ti, kk);
#pragma unroll
            for (int c = 0; c < NFC; ++c) {
                if (wv + 4 * c == kb) {
#pragma unroll
                    for (int ti = 0; ti < NT; ++ti) acc[c][ti] = panel_tile(a[ti], ti);
                } else {
#pragma unroll
                    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                        for (int kk = 0; kk < 4; ++kk) acc[c][ti] = Mfma<T>::run(a[ti][kk], b[c][kk], acc[c][ti]);
                }
            }
        }
#pragma unroll
        for (int x = 0; x < NX; ++x) {
            const int e = wv + 4 * x, eti = e % NT, etj = 4 * NFC + e / NT;
            if (e < NXT) {
                T ax[4];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) ax[kk] = ldA(eti, kk);
                if (etj == kb) {
                    accx[x] = panel_tile(ax, eti);
                } else {
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) accx[x] = Mfma<T>::run(ax[kk], bx[x][kk], accx[x]);
                }
            }
        }
        __syncthreads();
    }
    if (__builtin_amdgcn_readfirstlane(miss_sh)) {
        // another pivot than the diagonal was needed: M is untouched and the
        // block is left to k_fac_pivot (partial pivoting)
        if (tid == 0) redo[blk] = 1;
        return;
    }
    // the lane index again, opaque to the compiler: the store offsets are then
    // recomputed here instead of being kept live (spilled) since the loads
    int lane2;
    asm volatile("v_mov_b32 %0, %1" : "=v"(lane2) : "v"(lane0));
    auto stM = [&](int row, int col, T v) {
        if (row < n && col < n)
            *reinterpret_cast<T *>(reinterpret_cast<char *>(M) + (uint32_t)((col * n + row) * (int)sizeof(T))) = v;
    };
#pragma unroll
    for (int c = 0; c < NFC; ++c)
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                stM(16 * ti + Mfma<T>::row(lane2, r), 16 * (wv0 + 4 * c) + (lane2 & 15), acc[c][ti][r]);
#pragma unroll
    for (int x = 0; x < NX; ++x) {
        const int e = wv0 + 4 * x, eti = e % NT, etj = 4 * NFC + e / NT;
        if (e < NXT) {
#pragma unroll
            for (int r = 0; r < 4; ++r) stM(16 * eti + Mfma<T>::row(lane2, r), 16 * etj + (lane2 & 15), accx[x][r]);
        }
    }
}

// Partial-pivoting inverse of the blocks k_fac_binv flagged in redo[]; the
// other workgroups leave at once.
template <typename T, int NR>
__global__ void __launch_bounds__(kInvThreads) k_fac_pivot(T *__restrict__ D, int64_t ld, T *__restrict__ Wk, int64_t ldw,
                                                    const int32_t *__restrict__ rows, int n,
                                                    int32_t *__restrict__ err, int32_t *__restrict__ redo) {
    const int blk = rows[blockIdx.x];
    if (__builtin_amdgcn_readfirstlane(redo[blk]) == 0) return;
    extern __shared__ unsigned char smem_raw[];
    const int NP = 16 * ((n + 15) / 16), XPP = NP + 1;
    T *G = reinterpret_cast<T *>(smem_raw);
    T *XP = G + (size_t)NP * 17;
    T *prow_sh = XP + (size_t)16 * XPP;
    int32_t *rowtag = reinterpret_cast<int32_t *>(prow_sh + 16);
    int32_t *pivrow = rowtag + NP;
    __shared__ int32_t flag_sh;
    bool bad = false;
    ginv_pivot_body<T, NR, kInvThreads>(D + (size_t)blk * ld, Wk + (size_t)blockIdx.x * ldw, n, G, XP, prow_sh,
                                        rowtag, pivrow, &flag_sh, &bad);
    if (threadIdx.x == 0) {
        redo[blk] = 0;
        if (bad) atomicMin(err, blk);
    }
}

// Dynamic LDS of k_fac_ginv (the wide panel's halves) and k_fac_pivot (one
// 16-column panel and its pivot rows); both kernels get the larger.
template <typename T, int NR>
size_t ginv_lds(int n) {
    const int NP = 16 * ((n + 15) / 16);
    const size_t pivot_form = sizeof(T) * ((size_t)NP * 17 + (size_t)16 * (NP + 1) + 16) + sizeof(int32_t) * 2 * (size_t)NP;
    const size_t wide = sizeof(T) * (size_t)GinvHalves<T>::H * NP * 17 + sizeof(int32_t) * 2 * (size_t)NP;
    return std::max(wide, pivot_form);
}

__device__ __forceinline__ cdg_bf16 to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return static_cast<cdg_bf16>(u >> 16);
}

template <typename S> __device__ __forceinline__ S narrow(double v);
template <> __device__ __forceinline__ float narrow<float>(double v) { return static_cast<float>(v); }
template <> __device__ __forceinline__ cdg_bf16 narrow<cdg_bf16>(double v) { return to_bf16(static_cast<float>(v)); }

// Four narrowed values to dst[0..3] in one store (dst 8- or 16-byte aligned).
__device__ __forceinline__ void store4(float *dst, double a, double b, double c, double d) {
    *reinterpret_cast<float4 *>(dst) = make_float4((float)a, (float)b, (float)c, (float)d);
}
__device__ __forceinline__ void store4(cdg_bf16 *dst, double a, double b, double c, double d) {
    const uint32_t lo = (uint32_t)narrow<cdg_bf16>(a) | ((uint32_t)narrow<cdg_bf16>(b) << 16);
    const uint32_t hi = (uint32_t)narrow<cdg_bf16>(c) | ((uint32_t)narrow<cdg_bf16>(d) << 16);
    *reinterpret_cast<uint2 *>(dst) = make_uint2(lo, hi);
}

// U_u = (S) A_{up_blk[u]}   (the apply's copy of A's upper blocks).  One
// workgroup per block: 16-byte loads of two doubles, four values per thread and
// step, four steps' loads before their stores (slots start 512-byte aligned in
// A and 128-byte aligned in U).
template <typename S>
__global__ void __launch_bounds__(256) k_copy_upper(const double *__restrict__ Ab, int64_t lda, S *__restrict__ U,
                                                    int64_t ld, const int32_t *__restrict__ up_blk, int nup, int nn) {
    const int B = blockDim.x;
    for (int u = blockIdx.x; u < nup; u += gridDim.x) {
        const double2 *src = reinterpret_cast<const double2 *>(Ab + (size_t)up_blk[u] * lda);
        S *dst = U + (size_t)u * ld;
        const int n4 = nn / 4;
        int q = threadIdx.x;
        for (; q + 3 * B < n4; q += 4 * B) {
            double2 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = src[2 * (q + i * B)];
                b[i] = src[2 * (q + i * B) + 1];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) store4(dst + 4 * (q + i * B), a[i].x, a[i].y, b[i].x, b[i].y);
        }
        for (; q < n4; q += B) {
            const double2 a = src[2 * q], b = src[2 * q + 1];
            store4(dst + 4 * q, a.x, a.y, b.x, b.y);
        }
        const double *srcd = reinterpret_cast<const double *>(src);
        for (int e = 4 * n4 + threadIdx.x; e < nn; e += B) dst[e] = narrow<S>(srcd[e]);
    }
}

__global__ void k_f32_to_bf16(const float *__restrict__ src, cdg_bf16 *__restrict__ dst, int64_t count) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x)
        dst[k] = to_bf16(src[k]);
}

template <typename T, int NR>
int launch_ginv(cdg_ctx *ctx, cdg_prec *M, T *D, int64_t ld, const int32_t *rows, int nrows_lvl, int n,
                int32_t *err) {
    const size_t lds = ginv_lds<T, NR>(n);
    // Both kernels may take the LDS of the largest block this template serves
    // (n <= 64 NR, at most 320): set once per process, thread-safe (request-loop
    // workers launch concurrently), whatever block size comes first.
    // tiles each wave loads per round of the rank-16H update (two rounds' tiles
    // are in registers at once); CDG_GINV_TILES = 2, 4 or 8 overrides 2
    static const int ub = [] {
        const char *e = std::getenv("CDG_GINV_TILES");
        const int v = e ? std::atoi(e) : 2;
        return v <= 2 ? 2 : v <= 4 ? 4 : 8;
    }();
    const void *fn = ub == 2 ? reinterpret_cast<const void *>(&k_fac_ginv<T, NR, 2>)
                   : ub == 4 ? reinterpret_cast<const void *>(&k_fac_ginv<T, NR, 4>)
                             : reinterpret_cast<const void *>(&k_fac_ginv<T, NR, 8>);
    static std::once_flag once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(once, [] {
        const int most = (int)ginv_lds<T, NR>(std::min(64 * NR, 320));
        const void *fns[] = {reinterpret_cast<const void *>(&k_fac_ginv<T, NR, 2>),
                             reinterpret_cast<const void *>(&k_fac_ginv<T, NR, 4>),
                             reinterpret_cast<const void *>(&k_fac_ginv<T, NR, 8>),
                             reinterpret_cast<const void *>(&k_fac_pivot<T, NR>)};
        for (const void *f : fns)
            if (attr_err == hipSuccess)
                attr_err = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, most);
    });
    if (attr_err != hipSuccess)
        return cdg_fail(ctx, CDG_ERR_HIP, "hipFuncSetAttribute (block inverse LDS): %s", hipGetErrorString(attr_err));
    {
        T *Dp = D, *Wp = static_cast<T *>(M->inv_w);
        int64_t ldp = ld, ldw = M->ld;
        const int32_t *rp = rows;
        int np = n;
        int32_t *redo = M->d_redo;
        void *args[] = {&Dp, &ldp, &Wp, &ldw, &rp, &np, &redo};
        CDG_HIP_TRY(hipLaunchKernel(fn, dim3(nrows_lvl), dim3(kWideThreads), args, lds, ctx->stream));
    }
    CDG_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_fac_pivot<T, NR>), dim3(nrows_lvl), dim3(kInvThreads), lds, ctx->stream, D, ld,
                       static_cast<T *>(M->inv_w), M->ld, rows, n, err, M->d_redo);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

// Block-inverse kernel: 0 register-resident Gauss-Jordan (n <= 192), 1 blocked
// in registers with checked diagonal pivots (n <= 144), 2 blocked in global
// memory (any n <= 320; larger blocks always take it).  CDG_INV overrides
// the default, the fastest measured at 16^3: 1 up to n = 144, 0 above.
int inv_mode_env() {
    static const int mode = [] { const char *e = std::getenv("CDG_INV"); return e ? std::atoi(e) : -1; }();
    return mode;
}
template <typename T> int inv_mode(int n) {
    const int m = inv_mode_env();
    if (m >= 0) return m;
    return n <= 144 ? 1 : 0;
}

template <typename T>
int launch_inv(cdg_ctx *ctx, cdg_prec *M, T *D, int64_t ld, const int32_t *rows, int nrows_lvl, int n, int32_t *err) {
    if (!M->inv_w) return cdg_fail(ctx, CDG_ERR_STATE, "inverse scratch missing");
    const int mode = inv_mode<T>(n);
    if (n > 192 || mode == 2) {
        if (n <= 64) return launch_ginv<T, 1>(ctx, M, D, ld, rows, nrows_lvl, n, err);
        if (n <= 192) return launch_ginv<T, 3>(ctx, M, D, ld, rows, nrows_lvl, n, err);
        if (n <= 256) return launch_ginv<T, 4>(ctx, M, D, ld, rows, nrows_lvl, n, err);
        return launch_ginv<T, 5>(ctx, M, D, ld, rows, nrows_lvl, n, err);
    }
    dim3 grid(nrows_lvl), block(256);
    if (mode == 1 && n <= 144) {
        if (n <= 48) hipLaunchKernelGGL((k_fac_binv<T, 3>), grid, block, 0, ctx->stream, D, ld, rows, n, err, M->d_redo);
        else if (n <= 96) hipLaunchKernelGGL((k_fac_binv<T, 6>), grid, block, 0, ctx->stream, D, ld, rows, n, err, M->d_redo);
        else hipLaunchKernelGGL((k_fac_binv<T, 9>), grid, block, 0, ctx->stream, D, ld, rows, n, err, M->d_redo);
        CDG_HIP_TRY(hipGetLastError());
        // blocks that needed another pivot than the diagonal
        const size_t lds = ginv_lds<T, 3>(n);
        if (n <= 64) hipLaunchKernelGGL((k_fac_pivot<T, 1>), grid, dim3(kInvThreads), lds, ctx->stream, D, ld,
                                        static_cast<T *>(M->inv_w), M->ld, rows, n, err, M->d_redo);
        else hipLaunchKernelGGL((k_fac_pivot<T, 3>), grid, dim3(kInvThreads), lds, ctx->stream, D, ld,
                                static_cast<T *>(M->inv_w), M->ld, rows, n, err, M->d_redo);
        CDG_HIP_TRY(hipGetLastError());
        return CDG_OK;
    }
    if (n <= 48) hipLaunchKernelGGL((k_fac_inv<T, 3>), grid, block, 0, ctx->stream, D, ld, rows, n, err);
    else if (n <= 96) hipLaunchKernelGGL((k_fac_inv<T, 6>), grid, block, 0, ctx->stream, D, ld, rows, n, err);
    else if (n <= 144) hipLaunchKernelGGL((k_fac_inv<T, 9>), grid, block, 0, ctx->stream, D, ld, rows, n, err);
    else if (n <= 192) hipLaunchKernelGGL((k_fac_inv<T, 12>), grid, block, 0, ctx->stream, D, ld, rows, n, err);
    else return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "block inverse for n=%d not built", n);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

// Tile edge of the factor GEMMs: 64 above n = 192 (320 = 5 x 64), else 48
// (n = 135: 3 x 48 = 144, where 64-edge tiles would pad to 192).
inline bool gemm_wide(int n) { return n > 192; }

template <typename T, int TM>
int fold_all_tm(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const T *D) {
    const int n = M->n, tiles = (n + TM - 1) / TM;
    const int u1 = M->nup;
    const dim3 blk(Geo<TM>::GT);
    if (std::is_same<T, float>::value && M->precision == CDG_FP32 && M->U)
        // the upper blocks' FP32 copies (slot u; the values the FP64 reads convert to): half the bytes
        hipLaunchKernelGGL((k_fac_p1<T, float, TM>), dim3(xcd_grid(u1, tiles * tiles)), blk, 0, ctx->stream,
                           static_cast<const float *>(M->U), M->ld, static_cast<T *>(M->W), D, M->ld, M->d_up_all, u1,
                           M->d_up_all, M->d_up_col, n, tiles);
    else
        hipLaunchKernelGGL((k_fac_p1<T, double, TM>), dim3(xcd_grid(u1, tiles * tiles)), blk, 0, ctx->stream,
                           A->blocks, A->ld, static_cast<T *>(M->W), D, M->ld, M->d_up_all, u1, M->d_up_blk,
                           M->d_up_col, n, tiles);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

template <typename T>
int fold_all(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, const T *D) {
    if (M->nup == 0) return CDG_OK;
    return gemm_wide(M->n) ? fold_all_tm<T, 64>(ctx, M, A, D) : fold_all_tm<T, 48>(ctx, M, A, D);
}

// One level's GEMMs: L for the level's lower entries, then the rows' D.
template <typename T, int TM>
int gemm_level(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, T *L, T *D, const float *Ub32, int q0, int q1, int r0,
               int r1) {
    const int n = M->n, tiles = (n + TM - 1) / TM;
    const dim3 blk(Geo<TM>::GT);
    if (q1 > q0) {
        hipLaunchKernelGGL((k_fac_p1<T, double, TM>), dim3(xcd_grid(q1 - q0, tiles * tiles)), blk, 0, ctx->stream,
                           A->blocks, A->ld, L, D, M->ld, M->d_lev_lo + q0, q1 - q0, M->d_lo_blk, M->d_lo_col, n,
                           tiles);
        CDG_HIP_TRY(hipGetLastError());
    }
    if (r1 > r0) {
        if (Ub32)   // FP32 factor: the upper blocks' FP32 copies (the same values the FP64 reads convert to)
            hipLaunchKernelGGL((k_fac_p2<T, float, TM>), dim3(xcd_grid(r1 - r0, tiles * tiles)), blk, 0,
                               ctx->stream, A->blocks, A->ld, Ub32, M->ld, L, D, M->ld, M->d_lev_rows + r0, r1 - r0,
                               M->d_lo_ptr, M->d_lo_tru, M->d_diag, n, tiles);
        else
            hipLaunchKernelGGL((k_fac_p2<T, double, TM>), dim3(xcd_grid(r1 - r0, tiles * tiles)), blk, 0,
                               ctx->stream, A->blocks, A->ld, A->blocks, A->ld, L, D, M->ld, M->d_lev_rows + r0,
                               r1 - r0, M->d_lo_ptr, M->d_lo_tr, M->d_diag, n, tiles);
        CDG_HIP_TRY(hipGetLastError());
    }
    return CDG_OK;
}

template <typename T>
int factor_levels(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, T *L, T *D, const float *Ub32) {
    const int n = M->n;
    const bool fold = M->apply_mode == CDG_APPLY_FOLDED && M->W && M->precision != CDG_BF16;
    for (int l = 0; l < M->nlev; ++l) {
        const int q0 = M->lev_lo_ptr[l], q1 = M->lev_lo_ptr[l + 1];
        const int r0 = M->lev_row_ptr[l], r1 = M->lev_row_ptr[l + 1];
        int e = gemm_wide(n) ? gemm_level<T, 64>(ctx, M, A, L, D, Ub32, q0, q1, r0, r1)
                             : gemm_level<T, 48>(ctx, M, A, L, D, Ub32, q0, q1, r0, r1);
        if (e) return e;
        if (r1 > r0) {
            e = launch_inv<T>(ctx, M, D, M->ld, M->d_lev_rows + r0, r1 - r0, n, M->d_err);
            if (e) return e;
        }
    }
    if (fold) {   // W = U Dinv for all upper blocks, one batched launch
        int e = fold_all<T>(ctx, M, A, D);
        if (e) return e;
        M->w_ready = 1;
    }
    return CDG_OK;
}


}  // namespace

int launch_ilu_factor(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A, int32_t *bad_row) {
    const int n = M->n;
    if (n > 320) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "ILU factor for block size %d not built (n <= 320)", n);
    if (!M->d_redo) {
        int e = dev_alloc(ctx, (void **)&M->d_redo, sizeof(int32_t) * std::max(1, M->nrows));
        if (e) return e;
        M->owned.push_back({M->d_redo, sizeof(int32_t) * std::max(1, M->nrows)});
        CDG_HIP_TRY(hipMemsetAsync(M->d_redo, 0, sizeof(int32_t) * std::max(1, M->nrows), ctx->stream));
    }
    if (!M->inv_w) {   // global working slots of the block inverse, one per row of a level
        int maxrows = 1;
        for (int l = 0; l < M->nlev; ++l) maxrows = std::max(maxrows, M->lev_row_ptr[l + 1] - M->lev_row_ptr[l]);
        const size_t bytes = (M->precision == CDG_FP64 ? 8 : 4) * (size_t)maxrows * M->ld;
        int e = dev_alloc(ctx, &M->inv_w, bytes);
        if (e) return e;
        M->owned.push_back({M->inv_w, bytes});
    }
    const int32_t big = 0x7fffffff;
    CDG_HIP_TRY(hipMemcpyAsync(M->d_err, &big, sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    const int nn = n * n;
    // the FP32/BF16 copies of A's upper blocks (read by the backward sweep; the
    // FP32 factor's phase 2 reads its U_q from them too)
    if (M->precision != CDG_FP64 && M->nup > 0) {
        const dim3 grid(std::min(M->nup, 16384));
        if (M->precision == CDG_FP32)
            hipLaunchKernelGGL(k_copy_upper<float>, grid, dim3(256), 0, ctx->stream, A->blocks, A->ld,
                               static_cast<float *>(M->U), M->ld, M->d_up_blk, M->nup, nn);
        else
            hipLaunchKernelGGL(k_copy_upper<cdg_bf16>, grid, dim3(256), 0, ctx->stream, A->blocks, A->ld,
                               static_cast<cdg_bf16 *>(M->U), M->ld, M->d_up_blk, M->nup, nn);
        CDG_HIP_TRY(hipGetLastError());
    }
    int e = CDG_OK;
    if (M->precision == CDG_FP64) {
        e = factor_levels<double>(ctx, M, A, static_cast<double *>(M->L), static_cast<double *>(M->Dinv), nullptr);
    } else if (M->precision == CDG_FP32) {
        e = factor_levels<float>(ctx, M, A, static_cast<float *>(M->L), static_cast<float *>(M->Dinv),
                                 static_cast<const float *>(M->U));
    } else {
        e = factor_levels<float>(ctx, M, A, M->L32, M->D32, nullptr);
    }
    if (e) return e;
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

// W_u = Dinv_{row(u)} U_u in the row-slice layout (stream schedule), one
// batched launch over the upper entries.  FP64: A's blocks and the FP64 Dinv;
// FP32: the upper blocks' FP32 copies and Dinv; BF16: A's FP64 blocks (their
// values rounded to FP32 at the LDS store) and the FP32 working Dinv, the
// product stored BF16.
template <int TM>
int launch_fold_rs_tm(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A) {
    const int n = M->n, tiles = (n + TM - 1) / TM, nup = M->nup;
    const dim3 grid(xcd_grid(nup, tiles * tiles)), blk(Geo<TM>::GT);
    const int w = M->rs_w, SL = (int)M->rs_SL;
    if (M->precision == CDG_FP64)
        hipLaunchKernelGGL((k_fold_rs<double, double, double, TM>), grid, blk, 0, ctx->stream,
                           static_cast<const double *>(M->Dinv), M->ld, A->blocks, A->ld, M->d_up_row, M->d_up_blk, nup,
                           static_cast<double *>(M->Wrs), M->rs_slot, w, SL, n, tiles);
    else if (M->precision == CDG_FP32)
        hipLaunchKernelGGL((k_fold_rs<float, float, float, TM>), grid, blk, 0, ctx->stream,
                           static_cast<const float *>(M->Dinv), M->ld, static_cast<const float *>(M->U), M->ld,
                           M->d_up_row, M->d_up_all, nup, static_cast<float *>(M->Wrs), M->rs_slot, w, SL, n, tiles);
    else
        hipLaunchKernelGGL((k_fold_rs<float, double, cdg_bf16, TM>), grid, blk, 0, ctx->stream, M->D32, M->ld,
                           A->blocks, A->ld, M->d_up_row, M->d_up_blk, nup, static_cast<cdg_bf16 *>(M->Wrs), M->rs_slot,
                           w, SL, n, tiles);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int launch_fold_rs(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A) {
    if (M->nup == 0) return CDG_OK;
    return gemm_wide(M->n) ? launch_fold_rs_tm<64>(ctx, M, A) : launch_fold_rs_tm<48>(ctx, M, A);
}

int launch_fold_all(cdg_ctx *ctx, cdg_prec *M, const cdg_mat *A) {
    if (!M->W) return cdg_fail(ctx, CDG_ERR_STATE, "folded schedule without W storage");
    if (M->precision == CDG_FP64) return fold_all<double>(ctx, M, A, static_cast<const double *>(M->Dinv));
    if (M->precision == CDG_FP32) return fold_all<float>(ctx, M, A, static_cast<const float *>(M->Dinv));
    return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "folded schedule: FP64 and FP32 factors only");
}
