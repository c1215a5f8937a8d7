// Device-side synthetic Jacobian / vector generator (bench and test input; no
// reference counterpart).  Each value comes from integer hashing of
// (seed, block, element) plus single IEEE roundings with contraction disabled,
// so the result is bit-identical to the host generator used by the checker.
#include "cdg_internal.h"

#include <algorithm>
#include <vector>

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Irwin-Hall sum of four 32-bit uniforms recentred to zero mean, unit variance.
__device__ __forceinline__ double g_at(uint64_t seed, uint64_t idx) {
#pragma clang fp contract(off)
    const uint64_t G = 0x9E3779B97F4A7C15ULL;
    const uint64_t h1 = mix64(seed + G * (2 * idx + 1));
    const uint64_t h2 = mix64(seed + G * (2 * idx + 2));
    const uint64_t S = (h1 & 0xffffffffULL) + (h1 >> 32) + (h2 & 0xffffffffULL) + (h2 >> 32);
    const double u = (double)S * 2.3283064365386963e-10;
    return (u - 2.0) * 1.7320508075688772;
}

__global__ void k_fill_blocks(double *__restrict__ blocks, int64_t ld, const int8_t *__restrict__ isdiag, int nnzb,
                              int n, uint64_t seed, double s_diag, double s_off) {
#pragma clang fp contract(off)
    const int nn = n * n;
    for (int b = blockIdx.y; b < nnzb; b += gridDim.y) {
        const bool dg = isdiag[b] != 0;
        const double s = dg ? s_diag : s_off;
        double *B = blocks + (size_t)b * ld;
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += gridDim.x * blockDim.x) {
            const double g = g_at(seed, (uint64_t)b * 2097152ULL + (uint64_t)e);
            double val = g * s;
            const int c = e / n, r = e - c * n;
            if (dg && r == c) val = val + 2.0;
            B[e] = val;
        }
    }
}

__global__ void k_fill_vec(double *__restrict__ v, int64_t len, uint64_t seed, double scale) {
#pragma clang fp contract(off)
    const uint64_t sv = mix64(seed ^ 0xD1B54A32D192ED03ULL);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x)
        v[k] = g_at(sv, (uint64_t)k) * scale;
}

}  // namespace

int launch_fill_synthetic(cdg_ctx *ctx, cdg_mat *A, uint64_t seed, double s_diag, double s_off) {
    if (A->nnzb == 0) return CDG_OK;
    std::vector<int8_t> isdiag(A->nnzb, 0);
    for (int i = 0; i < A->nrows; ++i)
        for (int b = A->row_ptr[i]; b < A->row_ptr[i + 1]; ++b)
            isdiag[b] = (A->kind[b] == CDG_KIND_LOCAL && A->col[b] == i) ? 1 : 0;
    void *d = nullptr;
    int e = dev_alloc(ctx, &d, isdiag.size());
    if (e) return e;
    hipError_t he = hipMemcpyAsync(d, isdiag.data(), isdiag.size(), hipMemcpyHostToDevice, ctx->stream);
    if (he == hipSuccess) {
        const int nn = A->n * A->n;
        dim3 grid((unsigned)std::max(1, std::min((nn + 255) / 256, 64)), (unsigned)std::min(A->nnzb, 65535));
        hipLaunchKernelGGL(k_fill_blocks, grid, dim3(256), 0, ctx->stream, A->blocks, A->ld,
                           static_cast<const int8_t *>(d), A->nnzb, A->n, seed, s_diag, s_off);
        he = hipGetLastError();
        if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    }
    dev_free(ctx, d, isdiag.size());
    if (he != hipSuccess) return cdg_fail(ctx, CDG_ERR_HIP, "fill_synthetic: %s", hipGetErrorString(he));
    return CDG_OK;
}

int launch_vec_synthetic(cdg_ctx *ctx, int64_t len, uint64_t seed, double scale, double *v) {
    if (len == 0) return CDG_OK;
    dim3 grid((unsigned)std::min<int64_t>((len + 255) / 256, 4096));
    hipLaunchKernelGGL(k_fill_vec, grid, dim3(256), 0, ctx->stream, v, len, seed, scale);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}
