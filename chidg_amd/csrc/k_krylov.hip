// FP64 vector kernels of the FGMRES cycle (type_fgmres.f90:331-442 and the
// chidg_vector_t operators it uses).  Reductions are two-stage with a fixed
// order (per-workgroup partials, then one workgroup per result), so results
// are reproducible run to run.
#include "cdg_internal.h"

#include <algorithm>
#include <cstdlib>

namespace {

constexpr int RT = 256;   // threads per workgroup

// Workgroups of a reduction over len elements: ept elements per thread, at
// most kMaxParts workgroups (CDG_KRYLOV_EPT overrides ept).  Two elements (one
// 16-byte pair) per thread measured fastest for the FGMRES multidot and CGS
// update at 16^3 P2 (552,960 elements: 1,080 workgroups): the orthogonalization
// of a 38-iteration solve takes 3.3 ms, against 3.8 ms at 8 elements per
// thread (271 workgroups, about one per CU) and 3.6 ms at 1.
constexpr int kMaxParts = 2048;
int ept_setting() {
    static const int v = [] {
        const char *s = std::getenv("CDG_KRYLOV_EPT");
        const int x = s ? std::atoi(s) : 0;
        return x >= 1 && x <= 64 ? x : 2;
    }();
    return v;
}
int nparts(int64_t len) {
    const int64_t per = (int64_t)RT * ept_setting();
    return (int)std::min<int64_t>(kMaxParts, std::max<int64_t>(1, (len + per - 1) / per));
}

constexpr int NWAVE = RT / 64;

// Sum over the 64 lanes of a wave (butterfly: every lane gets the total).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup, returned to every thread: wave sums, then the
// NWAVE wave totals in wave order (sh holds NWAVE values).  A caller that
// reuses sh synchronizes before the next call.
__device__ __forceinline__ double block_sum(double v, double *sh) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = sh[0];
#pragma unroll
    for (int q = 1; q < NWAVE; ++q) s += sh[q];
    return s;
}

// part[i*nb + bx] = sum over chunk bx of V_i . w        (dot, operator_chidg_dot.f90:30-89)
// Workgroup (bx, g) takes vectors 8g..8g+7: w is read once per 8 products.
// With an even vector stride every thread reads element pairs (16-byte loads).
__global__ void __launch_bounds__(RT)
k_dot_part(int64_t len, const double *__restrict__ w, const double *__restrict__ V, int64_t vstride, int j,
           double *__restrict__ part, const double *__restrict__ gate) {
    if (gate && *gate == 0.0) return;   // a pass the host-side rule skips
    const int nb = gridDim.x, i0 = 8 * blockIdx.y;
    const int ni = min(8, j - i0);
    double acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = 0.0;
    const bool pairs = ((vstride | (int64_t)(reinterpret_cast<uintptr_t>(V) >> 3) |
                         (int64_t)(reinterpret_cast<uintptr_t>(w) >> 3)) & 1) == 0;
    if (pairs) {
        const int64_t len2 = len / 2;
        for (int64_t k2 = blockIdx.x * (int64_t)RT + threadIdx.x; k2 < len2; k2 += (int64_t)nb * RT) {
            const double2 wk = reinterpret_cast<const double2 *>(w)[k2];
            double2 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (u < ni) v[u] = reinterpret_cast<const double2 *>(V + (size_t)(i0 + u) * vstride)[k2];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (u < ni) {
                    acc[u] += v[u].x * wk.x;
                    acc[u] += v[u].y * wk.y;
                }
        }
        if (blockIdx.x == 0 && threadIdx.x == 0 && (len & 1)) {   // odd tail
            for (int u = 0; u < ni; ++u) acc[u] += V[(size_t)(i0 + u) * vstride + len - 1] * w[len - 1];
        }
    } else {
        for (int64_t k = blockIdx.x * (int64_t)RT + threadIdx.x; k < len; k += (int64_t)nb * RT) {
            const double wk = w[k];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (u < ni) acc[u] += V[(size_t)(i0 + u) * vstride + k] * wk;
        }
    }
    // the ni workgroup sums at once: wave sums, then thread u adds the wave totals of vector u
    __shared__ double shw[NWAVE][8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (u < ni) {
            const double t = wave_sum(acc[u]);
            if ((threadIdx.x & 63) == 0) shw[threadIdx.x >> 6][u] = t;
        }
    __syncthreads();
    if ((int)threadIdx.x < ni) {
        const int u = threadIdx.x;
        double t = shw[0][u];
#pragma unroll
        for (int q = 1; q < NWAVE; ++q) t += shw[q][u];
        part[(size_t)(i0 + u) * nb + blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(RT) k_sumsq_part(int64_t len, const double *__restrict__ w, double *__restrict__ part) {
    __shared__ double sh[NWAVE];
    double acc = 0.0;
    for (int64_t k = blockIdx.x * (int64_t)RT + threadIdx.x; k < len; k += (int64_t)gridDim.x * RT) acc += w[k] * w[k];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[i] = sum_b part[i*nb + b]
__global__ void __launch_bounds__(RT) k_reduce(const double *__restrict__ part, int nb, double *__restrict__ out,
                                               const double *__restrict__ gate = nullptr) {
    __shared__ double sh[NWAVE];
    if (gate && *gate == 0.0) return;
    const int i = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += RT) acc += part[(size_t)i * nb + b];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[i] = s;
}

// w = w - h_0 V_0 - h_1 V_1 - ... (sequential, as CGS's loop w = w - htmp(i)*v(i)),
// optional partial sums of the updated w^2 for the norm that follows.  The
// loads of 8 basis vectors go out before their products; with an even vector
// stride every thread updates element pairs (16-byte loads and stores).
__global__ void __launch_bounds__(RT)
k_cgs_update(int64_t len, double *__restrict__ w, const double *__restrict__ V, int64_t vstride, int j,
             const double *__restrict__ h, double *__restrict__ part, const double *__restrict__ gate) {
    __shared__ double sh[NWAVE];
    if (gate && *gate == 0.0) return;
    double acc = 0.0;
    const bool pairs = ((vstride | (int64_t)(reinterpret_cast<uintptr_t>(V) >> 3) |
                         (int64_t)(reinterpret_cast<uintptr_t>(w) >> 3)) & 1) == 0;
    if (pairs) {
        const int64_t len2 = len / 2;
        for (int64_t k2 = blockIdx.x * (int64_t)RT + threadIdx.x; k2 < len2; k2 += (int64_t)gridDim.x * RT) {
            double2 wk = reinterpret_cast<const double2 *>(w)[k2];
            for (int i0 = 0; i0 < j; i0 += 8) {
                const int ni = min(8, j - i0);
                double2 v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (u < ni) v[u] = reinterpret_cast<const double2 *>(V + (size_t)(i0 + u) * vstride)[k2];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (u < ni) {
                        wk.x = wk.x - h[i0 + u] * v[u].x;
                        wk.y = wk.y - h[i0 + u] * v[u].y;
                    }
            }
            reinterpret_cast<double2 *>(w)[k2] = wk;
            acc += wk.x * wk.x;
            acc += wk.y * wk.y;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0 && (len & 1)) {   // odd tail
            const int64_t k = len - 1;
            double wk = w[k];
            for (int i = 0; i < j; ++i) wk = wk - h[i] * V[(size_t)i * vstride + k];
            w[k] = wk;
            acc += wk * wk;
        }
    } else {
        for (int64_t k = blockIdx.x * (int64_t)RT + threadIdx.x; k < len; k += (int64_t)gridDim.x * RT) {
            double wk = w[k];
            for (int i = 0; i < j; ++i) wk = wk - h[i] * V[(size_t)i * vstride + k];
            w[k] = wk;
            acc += wk * wk;
        }
    }
    if (part) {
        const double s = block_sum(acc, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = s;
    }
}

// 1 when the selective re-orthogonalization rule asks for a second pass:
// sum_{i<jj} |h_i| / sqrt(w.w before) >= 1 (type_fgmres.f90:212-223), the
// same operations in the same order as the host's test.
__device__ __forceinline__ double reorth_flag(const double *h, int jj) {
    double crit = 0.0;
    for (int i = 0; i < jj; ++i) crit += fabs(h[i]);
    crit /= sqrt(h[jj]);
    return crit >= 1.0 ? 1.0 : 0.0;
}

__global__ void k_reorth_gate(const double *__restrict__ h, int jj, double *__restrict__ flag) {
    if (threadIdx.x != 0) return;
    *flag = reorth_flag(h, jj);
}

// k_reduce of one result, which also writes the re-orthogonalization flag of
// h[0..jj] (final before the update that produced the partials ran).
__global__ void __launch_bounds__(RT) k_reduce_gate(const double *__restrict__ part, int nb, double *__restrict__ out,
                                                    const double *__restrict__ h, int jj, double *__restrict__ flag) {
    __shared__ double sh[NWAVE];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += RT) acc += part[b];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        *out = s;
        *flag = reorth_flag(h, jj);
    }
}

__global__ void k_axpy_dev(int64_t len, double *__restrict__ w, const double *__restrict__ v, const double *__restrict__ h) {
    const double a = *h;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x)
        w[k] = w[k] - a * v[k];
}

__global__ void k_div(int64_t len, const double *__restrict__ w, double h, double *__restrict__ out) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x)
        out[k] = w[k] / h;
}

__global__ void k_div_inplace(int64_t len, double *__restrict__ w, double h) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x)
        w[k] = w[k] / h;
}

__global__ void k_sub(int64_t len, const double *__restrict__ b, const double *__restrict__ ax, double *__restrict__ r) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x)
        r[k] = b[k] - ax[k];
}

__global__ void k_add(int64_t len, double *__restrict__ z, const double *__restrict__ dz) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x)
        z[k] = z[k] + dz[k];
}

// x = x0 + y_0 Z_0 + y_1 Z_1 + ... (sequential, type_fgmres.f90:292-295)
__global__ void k_combine(int64_t len, double *__restrict__ x, const double *__restrict__ x0,
                          const double *__restrict__ Z, int64_t zstride, int kk, const double *__restrict__ y) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x) {
        double xk = x0[k];
        for (int i = 0; i < kk; ++i) xk = xk + y[i] * Z[(size_t)i * zstride + k];
        x[k] = xk;
    }
}

__global__ void k_axpby(int64_t len, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) y[i] = a * x[i] + b * y[i];
}

__global__ void k_fill(int64_t len, double *__restrict__ x, double v) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x)
        x[k] = v;
}

__global__ void k_copy(int64_t len, const double *__restrict__ s, double *__restrict__ d) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < len; k += (int64_t)gridDim.x * blockDim.x)
        d[k] = s[k];
}

dim3 egrid(int64_t len) { return dim3((unsigned)std::min<int64_t>(std::max<int64_t>(1, (len + 255) / 256), 2048)); }

}  // namespace

int kv_multidot(cdg_ctx *ctx, int64_t len, const double *w, const double *V, int64_t vstride, int j, double *d_out,
                const double *d_gate) {
    if (j <= 0) return CDG_OK;
    const int nb = nparts(len);
    int e = ctx_ensure_part(ctx, (size_t)nb * j);
    if (e) return e;
    hipLaunchKernelGGL(k_dot_part, dim3(nb, (j + 7) / 8), dim3(RT), 0, ctx->stream, len, w, V, vstride, j, ctx->d_part,
                       d_gate);
    hipLaunchKernelGGL(k_reduce, dim3(j), dim3(RT), 0, ctx->stream, ctx->d_part, nb, d_out, d_gate);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_sumsq(cdg_ctx *ctx, int64_t len, const double *w, double *d_out) {
    const int nb = nparts(len);
    int e = ctx_ensure_part(ctx, (size_t)nb);
    if (e) return e;
    hipLaunchKernelGGL(k_sumsq_part, dim3(nb), dim3(RT), 0, ctx->stream, len, w, ctx->d_part);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RT), 0, ctx->stream, ctx->d_part, nb, d_out);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_cgs_update(cdg_ctx *ctx, int64_t len, double *w, const double *V, int64_t vstride, int j, const double *d_h,
                  double *d_nrm2, const double *d_gate, double *d_flag) {
    const int nb = nparts(len);
    int e = ctx_ensure_part(ctx, (size_t)nb);
    if (e) return e;
    hipLaunchKernelGGL(k_cgs_update, dim3(nb), dim3(RT), 0, ctx->stream, len, w, V, vstride, j, d_h,
                       d_nrm2 ? ctx->d_part : nullptr, d_gate);
    if (d_nrm2 && d_flag && !d_gate) {
        hipLaunchKernelGGL(k_reduce_gate, dim3(1), dim3(RT), 0, ctx->stream, ctx->d_part, nb, d_nrm2, d_h, j, d_flag);
    } else {
        if (d_nrm2) hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RT), 0, ctx->stream, ctx->d_part, nb, d_nrm2, d_gate);
        if (d_flag) hipLaunchKernelGGL(k_reorth_gate, dim3(1), dim3(64), 0, ctx->stream, d_h, j, d_flag);
    }
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_reorth_gate(cdg_ctx *ctx, const double *d_h, int jj, double *d_flag) {
    hipLaunchKernelGGL(k_reorth_gate, dim3(1), dim3(64), 0, ctx->stream, d_h, jj, d_flag);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_axpy_dev(cdg_ctx *ctx, int64_t len, double *w, const double *v, const double *d_h) {
    hipLaunchKernelGGL(k_axpy_dev, egrid(len), dim3(256), 0, ctx->stream, len, w, v, d_h);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_div(cdg_ctx *ctx, int64_t len, const double *w, double h, double *out) {
    hipLaunchKernelGGL(k_div, egrid(len), dim3(256), 0, ctx->stream, len, w, h, out);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_div_inplace(cdg_ctx *ctx, int64_t len, double *w, double h) {
    hipLaunchKernelGGL(k_div_inplace, egrid(len), dim3(256), 0, ctx->stream, len, w, h);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_sub(cdg_ctx *ctx, int64_t len, const double *b, const double *ax, double *r) {
    hipLaunchKernelGGL(k_sub, egrid(len), dim3(256), 0, ctx->stream, len, b, ax, r);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_add(cdg_ctx *ctx, int64_t len, double *z, const double *dz) {
    hipLaunchKernelGGL(k_add, egrid(len), dim3(256), 0, ctx->stream, len, z, dz);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_combine(cdg_ctx *ctx, int64_t len, double *x, const double *x0, const double *Z, int64_t zstride, int k,
               const double *h_y) {
    int e = ctx_ensure_small(ctx, (size_t)k + 8);
    if (e) return e;
    double *d_y = ctx->d_small + 8;   // the first slots hold FGMRES scalars
    CDG_HIP_TRY(hipMemcpyAsync(d_y, h_y, sizeof(double) * (size_t)k, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_combine, egrid(len), dim3(256), 0, ctx->stream, len, x, x0, Z, zstride, k, d_y);
    CDG_HIP_TRY(hipGetLastError());
    // the host array may be reused right away
    CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return CDG_OK;
}

int kv_axpby(cdg_ctx *ctx, int64_t len, double a, const double *x, double b, double *y) {
    if (len == 0) return CDG_OK;
    hipLaunchKernelGGL(k_axpby, egrid(len), dim3(256), 0, ctx->stream, len, a, x, b, y);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_zero(cdg_ctx *ctx, int64_t len, double *x) {
    if (len == 0) return CDG_OK;
    hipLaunchKernelGGL(k_fill, egrid(len), dim3(256), 0, ctx->stream, len, x, 0.0);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

int kv_copy(cdg_ctx *ctx, int64_t len, const double *src, double *dst) {
    if (len == 0 || src == dst) return CDG_OK;
    hipLaunchKernelGGL(k_copy, egrid(len), dim3(256), 0, ctx->stream, len, src, dst);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}
