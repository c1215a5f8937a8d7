// Batched solve of many independent block-sparse systems (uneven sizes and
// block sizes, one to eight right-hand sides each) on one GPU: k runs of
// fgmres_t%solve (type_fgmres.f90:81-310, CGS, inner FGMRES off) per system,
// all advancing together.  No counterpart in the reference (ChiDG solves one
// system per MPI rank).
//
// * The systems of one block size are joined into one block-diagonal matrix
//   (rows concatenated, no coupling between systems).  Its natural-order
//   ILU(0) is each system's own factor (no row of one system touches another
//   system's rows), so one factor launch sequence serves all of them and each
//   wavefront level of the apply is one launch across the systems.
// * Lanes are the (system, right-hand side) pairs.  Right-hand side r of
//   every system sits in column r of one vector space (columns of Ltot
//   values, the systems one after another); the k-vector matvec and row-slice
//   apply of each joined matrix read every block once for the k right-hand
//   sides of its system (rows of a system with fewer right-hand sides leave
//   the other columns alone).
// * The Arnoldi step's vector work runs segmented: one launch covers every
//   lane's segment (work items of 2048 values), with per-lane reductions in a
//   fixed order; the rotations, the residual estimate and the stopping test
//   of every lane run on the device (one thread per lane).  Each Arnoldi
//   step is captured once as a HIP graph and replayed; the host never waits
//   for the step it launched, only for the one two steps back (its lanes'
//   active flags copied to pinned memory behind an event), and launches no
//   further step once every lane has stopped.  Lanes that stop mid-cycle are
//   skipped by the remaining steps (rows of fully stopped systems too).  The
//   host runs the cycle ends (back substitution of the small triangular
//   systems, restarts), as fgmres_t%solve does.
#include "cdg_internal.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

constexpr int BT = 256;          // threads per workgroup of the segmented kernels
constexpr int BPER = 8;          // values per thread of one work item
constexpr int BCH = BT * BPER;   // values per work item
constexpr int BMAXJ = 32;        // basis vectors per lane at most (nkrylov <= 31)
constexpr int BQ = BMAXJ + 1;    // partial sums per work item
constexpr int BD = 2 * BMAXJ + 8;// per-lane reduced values

__device__ __forceinline__ double bwave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct Seg {
    const int32_t *item_lane;    // lane of each work item
    const int32_t *item_start;   // first value of the item inside its lane's segment
    const int64_t *lane_off;     // offset of a lane's segment inside one basis step
    const int32_t *lane_len;
    const int32_t *lane_item0;   // first work item of each lane
    const int32_t *lane_nitem;
    int64_t kstride;             // values between consecutive basis vectors (K * Ltot)
    int32_t nlanes;
};

// part[it][i] = w . V_i over the item, i < jj; with sq also part[it][jj] = w . w.
// Lanes whose flag is 0 (and, with gate, whose gate is 0) are skipped.
__global__ void __launch_bounds__(BT) k_bdots(Seg s, const double *__restrict__ V, const double *__restrict__ W,
                                              int jj, int sq, const int32_t *__restrict__ flag,
                                              const double *__restrict__ gate, double *__restrict__ part) {
    const int it = blockIdx.x, l = s.item_lane[it];
    if (!flag[l] || (gate && gate[(size_t)l * BD] == 0.0)) return;
    const int64_t o = s.lane_off[l] + s.item_start[it];
    const int cnt = min(BCH, s.lane_len[l] - s.item_start[it]);
    double wv[BPER];
#pragma unroll
    for (int u = 0; u < BPER; ++u) {
        const int e = threadIdx.x + BT * u;
        wv[u] = e < cnt ? W[o + e] : 0.0;
    }
    double acc[BQ];
#pragma unroll
    for (int i = 0; i < BQ; ++i) acc[i] = 0.0;
#pragma unroll
    for (int i = 0; i < BMAXJ; ++i) {
        if (i < jj) {
            const double *vi = V + (size_t)i * s.kstride + o;
#pragma unroll
            for (int u = 0; u < BPER; ++u) {
                const int e = threadIdx.x + BT * u;
                if (e < cnt) acc[i] += vi[e] * wv[u];
            }
        }
    }
    if (sq)
#pragma unroll
        for (int u = 0; u < BPER; ++u) acc[BMAXJ] += wv[u] * wv[u];
    __shared__ double sh[BT / 64][BQ];
    const int nq = jj;
#pragma unroll
    for (int i = 0; i < BQ; ++i) {
        if (i < nq || (i == BMAXJ && sq)) {
            const double t = bwave_sum(acc[i]);
            if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][i] = t;
        }
    }
    __syncthreads();
    const int q = threadIdx.x;
    if (q < nq || (q == BMAXJ && sq)) {
        double t = sh[0][q];
#pragma unroll
        for (int wv2 = 1; wv2 < BT / 64; ++wv2) t += sh[wv2][q];
        part[(size_t)it * BQ + q] = t;
    }
}

// out[l][ooff + q] = sum over the lane's items of part[it][q], q < nq, items
// in order; q = BMAXJ (the w.w sum) goes to out[l][ooff + nq].  With flagout,
// the selective re-orthogonalization rule (type_fgmres.f90:212-223) of the
// first pass: flagout[l] = 1 when sum_{i<nq} |h_i| / sqrt(w.w) >= 1.
__global__ void __launch_bounds__(64) k_breduce(Seg s, const double *__restrict__ part, int nq, int sq,
                                                const int32_t *__restrict__ flag, const double *__restrict__ gate,
                                                double *__restrict__ out, int ooff, double *__restrict__ flagout) {
    const int l = blockIdx.x;
    if (!flag[l] || (gate && gate[(size_t)l * BD] == 0.0)) return;
    const int it0 = s.lane_item0[l], nit = s.lane_nitem[l];
    double *o = out + (size_t)l * BD + ooff;
    for (int q = threadIdx.x; q < nq + sq; q += 64) {
        const int src = q < nq ? q : BMAXJ;
        double t = 0.0;
        for (int it = 0; it < nit; ++it) t += part[(size_t)(it0 + it) * BQ + src];
        o[q] = t;
    }
    if (flagout) {
        __syncthreads();
        if (threadIdx.x == 0) {
            double crit = 0.0;
            for (int i = 0; i < nq; ++i) crit += fabs(o[i]);
            crit /= sqrt(o[nq]);
            flagout[(size_t)l * BD] = crit >= 1.0 ? 1.0 : 0.0;
        }
    }
}

// w -= sum_i h_i V_i (i in order, classical_gram_schmidt's update), and
// part[it][BMAXJ] = w . w over the item.
__global__ void __launch_bounds__(BT) k_bupdate(Seg s, const double *__restrict__ V, double *__restrict__ W, int jj,
                                                const double *__restrict__ h, const int32_t *__restrict__ flag,
                                                const double *__restrict__ gate, double *__restrict__ part) {
    const int it = blockIdx.x, l = s.item_lane[it];
    if (!flag[l] || (gate && gate[(size_t)l * BD] == 0.0)) return;
    const int64_t o = s.lane_off[l] + s.item_start[it];
    const int cnt = min(BCH, s.lane_len[l] - s.item_start[it]);
    const double *hl = h + (size_t)l * BD;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < BPER; ++u) {
        const int e = threadIdx.x + BT * u;
        if (e < cnt) {
            double wk = W[o + e];
            for (int i = 0; i < jj; ++i) wk = wk - hl[i] * V[(size_t)i * s.kstride + o + e];
            W[o + e] = wk;
            acc += wk * wk;
        }
    }
    __shared__ double sh[BT / 64];
    const double t = bwave_sum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int q = 1; q < BT / 64; ++q) r += sh[q];
        part[(size_t)it * BQ + BMAXJ] = r;
    }
}

// dst = src / f[l] over the lanes whose flag is set
__global__ void __launch_bounds__(BT) k_bdiv(Seg s, const double *__restrict__ src, double *__restrict__ dst,
                                             const double *__restrict__ f, int fstride, const int32_t *__restrict__ flag) {
    const int it = blockIdx.x, l = s.item_lane[it];
    if (!flag[l]) return;
    const int64_t o = s.lane_off[l] + s.item_start[it];
    const int cnt = min(BCH, s.lane_len[l] - s.item_start[it]);
    const double d = f[(size_t)l * fstride];
#pragma unroll
    for (int u = 0; u < BPER; ++u) {
        const int e = threadIdx.x + BT * u;
        if (e < cnt) dst[o + e] = src[o + e] / d;
    }
}

// x = x0 + sum_{i < nv[l]} y[l][i] Z_i (type_fgmres.f90:287-294), then, for
// lanes that restart, x0 = x
__global__ void __launch_bounds__(BT) k_bcombine(Seg s, const double *__restrict__ Z, double *__restrict__ X,
                                                 double *__restrict__ X0, const double *__restrict__ y,
                                                 const int32_t *__restrict__ nv, const int32_t *__restrict__ restart) {
    const int it = blockIdx.x, l = s.item_lane[it];
    const int k = nv[l];
    if (k <= 0) return;
    const int64_t o = s.lane_off[l] + s.item_start[it];
    const int cnt = min(BCH, s.lane_len[l] - s.item_start[it]);
    const double *yl = y + (size_t)l * BMAXJ;
#pragma unroll
    for (int u = 0; u < BPER; ++u) {
        const int e = threadIdx.x + BT * u;
        if (e < cnt) {
            double v = X0[o + e];
            for (int i = 0; i < k; ++i) v += yl[i] * Z[(size_t)i * s.kstride + o + e];
            X[o + e] = v;
            if (restart[l]) X0[o + e] = v;
        }
    }
}

// row_k[i] = 1 + the last right-hand side of row i's system whose lane is
// still active (0 when none is): the step's applies and matvecs skip the
// rows of systems whose lanes have all stopped.
__global__ void k_browk(int nrows, const int32_t *__restrict__ row_sys, const int32_t *__restrict__ sys_lane0,
                        const int32_t *__restrict__ sys_nrhs, const int32_t *__restrict__ active,
                        int32_t *__restrict__ row_k) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const int sy = row_sys[i], l0 = sys_lane0[sy];
    int k = 0;
    for (int r = 0; r < sys_nrhs[sy]; ++r)
        if (active[l0 + r]) k = r + 1;
    row_k[i] = k;
}

__global__ void k_bsub(int64_t len, const double *__restrict__ b, double *__restrict__ r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) r[i] = b[i] - r[i];
}

struct LaneState {
    int32_t *active;    // taking Arnoldi steps in this cycle
    int32_t *stepped;   // took step j (set by the rotation kernel)
    int32_t *nvecs, *niter, *conv, *nreorth, *err;
    double *H;          // [lane][(m+1) m], column-major
    double *cs, *sn, *p;// [lane][m+1]
    double *res, *r0init, *hn;
};

// One Arnoldi step's rotations per lane (type_fgmres.f90:229-280): column j
// of H from the two CGS passes, h(j+1,j) = ||w||, the previous rotations, the
// new one, p, the residual estimate and the stopping test.
__global__ void k_bgivens(int nlanes, int j, int m, LaneState st, const double *__restrict__ d1,
                          const double *__restrict__ d2, double tol, double rtol, int nmax) {
#pragma clang fp contract(off)
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlanes) return;
    if (!st.active[l]) { st.stepped[l] = 0; return; }
    st.stepped[l] = 1;
    const int jj = j + 1, M1 = m + 1;
    double *H = st.H + (size_t)l * M1 * m;
    double *c = st.cs + (size_t)l * M1, *s = st.sn + (size_t)l * M1, *p = st.p + (size_t)l * M1;
    const double *a = d1 + (size_t)l * BD, *b = d2 + (size_t)l * BD;
    const bool re = a[jj + 2] != 0.0;   // the first pass's flag (second pass ran)
    for (int i = 0; i < jj; ++i) {
        double h = 0.0;
        h += a[i];
        if (re) h += b[i];
        H[(size_t)j * M1 + i] = h;
    }
    const double nrm2 = re ? b[jj] : a[jj + 1];
    if (re) st.nreorth[l] += 1;
    H[(size_t)j * M1 + j + 1] = sqrt(nrm2);
    st.hn[l] = H[(size_t)j * M1 + j + 1];
    for (int i = 0; i < j; ++i) {
        const double hij = c[i] * H[(size_t)j * M1 + i] + s[i] * H[(size_t)j * M1 + i + 1];
        const double hipj = -s[i] * H[(size_t)j * M1 + i] + c[i] * H[(size_t)j * M1 + i + 1];
        H[(size_t)j * M1 + i] = hij;
        H[(size_t)j * M1 + i + 1] = hipj;
    }
    const double gam = hypot(H[(size_t)j * M1 + j], H[(size_t)j * M1 + j + 1]);
    c[j] = H[(size_t)j * M1 + j] / gam;
    s[j] = H[(size_t)j * M1 + j + 1] / gam;
    H[(size_t)j * M1 + j] = gam;
    H[(size_t)j * M1 + j + 1] = 0.0;
    const double pj = c[j] * p[j], pjp = -s[j] * p[j];
    p[j] = pj;
    p[j + 1] = pjp;
    st.niter[l] += 1;
    st.nvecs[l] = jj;
    const double res = fabs(p[j + 1]);
    st.res[l] = res;
    if (!isfinite(res)) { st.err[l] = 1; st.active[l] = 0; return; }
    const bool ca = res < tol, cr = (res / st.r0init[l]) < rtol, cn = (st.niter[l] >= nmax) && (nmax > 0);
    if (ca || cr || cn) { st.conv[l] = 1; st.active[l] = 0; }
}

}  // namespace

struct BatchGroup {
    int n = 0;
    std::vector<int> members;          // system indices
    cdg_mat *U = nullptr;              // joined matrix
    cdg_prec *M = nullptr;             // its ILU(0)
    int32_t *d_row_k = nullptr;        // right-hand sides per joined row (the system's nrhs)
    int32_t *d_row_kl = nullptr;       // the same, limited to the lanes still active (per step)
    int32_t *d_row_sys = nullptr;      // system of each joined row
    int64_t goff = 0;                  // first value of the group in a column
    // row ranges of systems with like right-hand-side counts (members are in
    // ascending nrhs): {row_lo, row_hi, vectors per launch}
    std::vector<std::array<int, 3>> mv_ranges, ap_ranges;
};

struct cdg_batch {
    int nsys = 0, K = 1, m = 30, precision = CDG_FP32;
    int64_t Ltot = 0;
    std::vector<int64_t> off, len;     // per system: offset in a column, length
    std::vector<int> nrhs;
    std::vector<int64_t> sys_nnzb;     // per system: blocks and pattern (for cdg_batch_update's checks)
    std::vector<int> sys_nrows, sys_n;
    std::vector<std::vector<int32_t>> sys_rp, sys_col, sys_kind;
    std::vector<BatchGroup> groups;
    // lanes and work items
    int nlanes = 0, nitems = 0;
    std::vector<int> lane_sys, lane_r;
    std::vector<int32_t> h_lane_item0, h_lane_nitem;
    Seg seg{};
    // device memory
    std::vector<DevMem> owned;
    double *V = nullptr, *Z = nullptr, *B = nullptr, *X = nullptr, *X0 = nullptr, *R0 = nullptr;
    void *Y = nullptr;                 // forward vectors of the applies (K * Ltot values of 8 or 4 bytes)
    double *part = nullptr, *d1 = nullptr, *d2 = nullptr, *ycomb = nullptr, *scale0 = nullptr;
    int32_t *live = nullptr, *nv = nullptr, *restart = nullptr;
    LaneState st{};
    int32_t *d_sys_lane0 = nullptr, *d_sys_nrhs = nullptr;
    std::vector<hipGraph_t> graphs;    // one captured graph per Arnoldi step
    std::vector<hipGraphExec_t> execs;
    double graph_tol = -1.0, graph_rtol = -1.0;   // the stopping test the graphs were captured with
    int graph_nmax = 0;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int32_t *h_act = nullptr;          // pinned: three copies of the lanes' active flags (one per event)
};

namespace {

int balloc(cdg_ctx *ctx, cdg_batch *bt, void **p, size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    int e = dev_alloc(ctx, p, bytes);
    if (e) return e;
    bt->owned.push_back({*p, bytes});
    CDG_HIP_TRY(hipMemsetAsync(*p, 0, bytes, ctx->stream));
    return CDG_OK;
}

template <typename T>
int bupload(cdg_ctx *ctx, cdg_batch *bt, T **d, const std::vector<T> &h) {
    int e = balloc(ctx, bt, (void **)d, sizeof(T) * std::max<size_t>(h.size(), 1));
    if (e) return e;
    if (!h.empty()) CDG_HIP_TRY(hipMemcpyAsync(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    return CDG_OK;
}

// Y = A X over a group's rows, each range of like systems with its own vector count
int group_mv(cdg_ctx *ctx, cdg_batch *bt, const BatchGroup &g, const double *X, double *Y, const int32_t *row_k) {
    for (const auto &r : g.mv_ranges) {
        int e = launch_mv_multi(ctx, g.U, X + g.goff, Y + g.goff, bt->Ltot, r[2], row_k, r[0], r[1]);
        if (e) return e;
    }
    return CDG_OK;
}

// The launches of Arnoldi step j of every lane (no host synchronization).
int batch_step(cdg_ctx *ctx, cdg_batch *bt, int j, double tol, double rtol, int nmax) {
    const int m = bt->m, jj = j + 1;
    const int64_t ks = bt->seg.kstride;
    double *Vj = bt->V + (size_t)j * ks, *Zj = bt->Z + (size_t)j * ks, *W = bt->V + (size_t)(j + 1) * ks;
    const size_t yv = bt->precision == CDG_FP64 ? 8 : 4;   // forward vectors hold the accumulation type
    for (BatchGroup &g : bt->groups) {   // z_j = M^-1 v_j, w = A z_j: each block read once per system
        hipLaunchKernelGGL(k_browk, dim3((g.U->nrows + 255) / 256), dim3(256), 0, ctx->stream, g.U->nrows, g.d_row_sys,
                           bt->d_sys_lane0, bt->d_sys_nrhs, bt->st.active, g.d_row_kl);
        for (const auto &r : g.ap_ranges) {
            int e = launch_ilu_apply_rs_multi(ctx, g.M, Vj + g.goff, Zj + g.goff,
                                              static_cast<char *>(bt->Y) + yv * g.goff, bt->Ltot, r[2], g.d_row_kl,
                                              r[0], r[1]);
            if (e) return e;
        }
        int e = group_mv(ctx, bt, g, Zj, W, g.d_row_kl);
        if (e) return e;
    }
    const Seg &s = bt->seg;
    const dim3 gi(bt->nitems), gl(bt->nlanes);
    int32_t *act = bt->st.active;
    // first CGS pass: products and w.w, the re-orthogonalization flag, the update and its w.w
    hipLaunchKernelGGL(k_bdots, gi, dim3(BT), 0, ctx->stream, s, bt->V, W, jj, 1, act, nullptr, bt->part);
    hipLaunchKernelGGL(k_breduce, gl, dim3(64), 0, ctx->stream, s, bt->part, jj, 1, act, nullptr, bt->d1, 0,
                       bt->d1 + jj + 2);
    hipLaunchKernelGGL(k_bupdate, gi, dim3(BT), 0, ctx->stream, s, bt->V, W, jj, bt->d1, act, nullptr, bt->part);
    hipLaunchKernelGGL(k_breduce, gl, dim3(64), 0, ctx->stream, s, bt->part, 0, 1, act, nullptr, bt->d1, jj + 1,
                       nullptr);
    // gated second pass
    const double *gate = bt->d1 + jj + 2;
    hipLaunchKernelGGL(k_bdots, gi, dim3(BT), 0, ctx->stream, s, bt->V, W, jj, 0, act, gate, bt->part);
    hipLaunchKernelGGL(k_breduce, gl, dim3(64), 0, ctx->stream, s, bt->part, jj, 0, act, gate, bt->d2, 0, nullptr);
    hipLaunchKernelGGL(k_bupdate, gi, dim3(BT), 0, ctx->stream, s, bt->V, W, jj, bt->d2, act, gate, bt->part);
    hipLaunchKernelGGL(k_breduce, gl, dim3(64), 0, ctx->stream, s, bt->part, 0, 1, act, gate, bt->d2, jj, nullptr);
    // rotations and the stopping test; then v_{j+1} = w / h(j+1,j) for the lanes that took the step
    hipLaunchKernelGGL(k_bgivens, dim3((bt->nlanes + 63) / 64), dim3(64), 0, ctx->stream, bt->nlanes, j, m, bt->st,
                       bt->d1, bt->d2, tol, rtol, nmax);
    hipLaunchKernelGGL(k_bdiv, gi, dim3(BT), 0, ctx->stream, s, W, W, bt->st.hn, 1, bt->st.stepped);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}

using bclk = std::chrono::steady_clock;

// a HIP wait inside cdg_batch_create: the error goes to e (the batch is then destroyed)
int bwait(cdg_ctx *ctx, hipError_t he, const char *what) {
    return he == hipSuccess ? CDG_OK : cdg_fail(ctx, CDG_ERR_HIP, "batch: %s: %s", what, hipGetErrorString(he));
}

void batch_drop_graphs(cdg_batch *bt) {
    for (hipGraphExec_t x : bt->execs)
        if (x) (void)hipGraphExecDestroy(x);
    for (hipGraph_t g : bt->graphs)
        if (g) (void)hipGraphDestroy(g);
    bt->execs.clear();
    bt->graphs.clear();
}

// CDG_BATCH_TRACE=1: phase times of create and solve on stderr
bool btrace() {
    static const bool on = std::getenv("CDG_BATCH_TRACE") && std::atoi(std::getenv("CDG_BATCH_TRACE")) != 0;
    return on;
}

double bsecs(bclk::time_point a) { return std::chrono::duration<double>(bclk::now() - a).count(); }

}  // namespace

extern "C" int cdg_batch_create(cdg_ctx *ctx, int32_t nsys, cdg_mat *const *A, const int32_t *nrhs, int32_t precision,
                                int32_t nkrylov, cdg_batch **out) {
    if (!ctx || !A || !nrhs || !out || nsys < 1) return CDG_ERR_ARG;
    if (nkrylov < 1 || nkrylov >= BMAXJ) return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "batch: nkrylov must be 1..%d", BMAXJ - 1);
    if (precision != CDG_FP64 && precision != CDG_FP32 && precision != CDG_BF16)
        return cdg_fail(ctx, CDG_ERR_ARG, "unknown precision %d", precision);
    *out = nullptr;
    const auto t0 = bclk::now();
    cdg_batch *bt = new cdg_batch();
    bt->nsys = nsys;
    bt->m = nkrylov;
    bt->precision = precision;
    bt->nrhs.assign(nrhs, nrhs + nsys);
    for (int s = 0; s < nsys; ++s) {
        if (!A[s] || nrhs[s] < 1 || nrhs[s] > 8) { delete bt; return cdg_fail(ctx, CDG_ERR_ARG, "batch: system %d", s); }
        if (!A[s]->part.empty()) {
            delete bt;
            return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "batch: system %d has a rank partition", s);
        }
        bt->K = std::max(bt->K, (int)nrhs[s]);
        bt->sys_nnzb.push_back(A[s]->nnzb);
        bt->sys_nrows.push_back(A[s]->nrows);
        bt->sys_n.push_back(A[s]->n);
        bt->sys_rp.push_back(A[s]->row_ptr);
        bt->sys_col.push_back(A[s]->col);
        bt->sys_kind.push_back(A[s]->kind);
    }
    // groups by block size, systems of a group one after another in a column
    std::vector<int> order(nsys);
    for (int s = 0; s < nsys; ++s) order[s] = s;
    // by block size, then by right-hand sides (like systems in one row range of the joined matrix)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return A[a]->n != A[b]->n ? A[a]->n < A[b]->n : nrhs[a] < nrhs[b];
    });
    bt->off.assign(nsys, 0);
    bt->len.assign(nsys, 0);
    int64_t pos = 0;
    for (int idx = 0; idx < nsys; ++idx) {
        const int s = order[idx];
        if (bt->groups.empty() || bt->groups.back().n != A[s]->n) {
            bt->groups.emplace_back();
            bt->groups.back().n = A[s]->n;
            bt->groups.back().goff = pos;
        }
        bt->groups.back().members.push_back(s);
        bt->off[s] = pos;
        bt->len[s] = (int64_t)A[s]->nrows * A[s]->n;
        pos += bt->len[s];
    }
    bt->Ltot = pos;
    int e = bwait(ctx, hipDeviceSynchronize(), "synchronize");   // the callers' matrices are complete
    // joined matrices: concatenated patterns, blocks copied device to device
    for (BatchGroup &g : bt->groups) {
        std::vector<int32_t> rp(1, 0), col, kind, rowk, rowsys;
        int nrows = 0;
        for (int s : g.members) {
            const cdg_mat *a = A[s];
            for (int i = 0; i < a->nrows; ++i) {
                for (int b = a->row_ptr[i]; b < a->row_ptr[i + 1]; ++b) {
                    col.push_back(a->col[b] + nrows);
                    kind.push_back(a->kind[b]);
                }
                rp.push_back((int32_t)col.size());
                rowk.push_back(bt->nrhs[s]);
                rowsys.push_back(s);
            }
            nrows += a->nrows;
        }
        const auto tm = bclk::now();
        if (!e) e = cdg_mat_create(ctx, nrows, g.n, rp.data(), col.data(), kind.data(), &g.U);
        const double t_mat = bsecs(tm);
        size_t b0 = 0;
        for (int s : g.members) {
            if (e) break;
            const size_t bytes = sizeof(double) * (size_t)A[s]->nnzb * A[s]->ld;
            if (hipMemcpyAsync(g.U->blocks + b0 * g.U->ld, A[s]->blocks, bytes, hipMemcpyDeviceToDevice, ctx->stream) !=
                hipSuccess)
                e = cdg_fail(ctx, CDG_ERR_HIP, "batch: copying the blocks of system %d", s);
            b0 += A[s]->nnzb;
        }
        {   // row ranges by right-hand-side class: matvecs with 1, 2, 4, 8 vectors per launch,
            // applies with 2, 4, 8 (the row-slice multi kernel's widths)
            auto cls = [](int k) { return k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : 8; };
            int row = 0;
            for (int s : g.members) {
                const int lo = row, hi = row + A[s]->nrows, km = cls(bt->nrhs[s]), ka = std::max(2, km);
                row = hi;
                if (!g.mv_ranges.empty() && g.mv_ranges.back()[2] == km) g.mv_ranges.back()[1] = hi;
                else g.mv_ranges.push_back({lo, hi, km});
                if (!g.ap_ranges.empty() && g.ap_ranges.back()[2] == ka) g.ap_ranges.back()[1] = hi;
                else g.ap_ranges.push_back({lo, hi, ka});
            }
        }
        if (!e) e = bupload(ctx, bt, &g.d_row_k, rowk);
        if (!e) e = bupload(ctx, bt, &g.d_row_kl, rowk);
        if (!e) e = bupload(ctx, bt, &g.d_row_sys, rowsys);
        if (!e) e = bwait(ctx, hipStreamSynchronize(ctx->stream), "block copies");
        const double t_copy = bsecs(tm) - t_mat;
        if (!e) e = cdg_prec_create(ctx, g.U, CDG_PREC_ILU0, precision, &g.M);
        if (!e) e = cdg_prec_set_schedule(ctx, g.M, CDG_SCHEDULE_ROWSLICE);
        if (btrace())
            std::fprintf(stderr, "[batch] join n=%d: matrix %.3f s, block copies %.3f s, preconditioner %.3f s\n", g.n,
                         t_mat, t_copy, bsecs(tm) - t_mat - t_copy);
    }
    if (!e) e = bwait(ctx, hipStreamSynchronize(ctx->stream), "join");
    const double t_join = bsecs(t0);
    // lanes and work items
    std::vector<int64_t> lane_off;
    std::vector<int32_t> lane_len, item_lane, item_start;
    for (int s = 0; s < nsys && !e; ++s)
        for (int r = 0; r < bt->nrhs[s]; ++r) {
            const int l = bt->nlanes++;
            bt->lane_sys.push_back(s);
            bt->lane_r.push_back(r);
            lane_off.push_back((int64_t)r * bt->Ltot + bt->off[s]);
            lane_len.push_back((int32_t)bt->len[s]);
            bt->h_lane_item0.push_back((int32_t)item_lane.size());
            int cnt = 0;
            for (int64_t st = 0; st < bt->len[s]; st += BCH, ++cnt) {
                item_lane.push_back(l);
                item_start.push_back((int32_t)st);
            }
            bt->h_lane_nitem.push_back(cnt);
        }
    bt->nitems = (int)item_lane.size();
    {   // first lane and right-hand sides of each system (lanes are numbered system by system)
        std::vector<int32_t> l0(nsys, 0);
        for (int s = 1; s < nsys; ++s) l0[s] = l0[s - 1] + bt->nrhs[s - 1];
        if (!e) e = bupload(ctx, bt, &bt->d_sys_lane0, l0);
        if (!e) e = bupload(ctx, bt, &bt->d_sys_nrhs, bt->nrhs);
    }
    const size_t col = (size_t)bt->K * bt->Ltot, M1 = (size_t)bt->m + 1, nl = bt->nlanes;
    int32_t *d_item_lane = nullptr, *d_item_start = nullptr, *d_lane_len = nullptr, *d_item0 = nullptr, *d_nitem = nullptr;
    int64_t *d_lane_off = nullptr;
    if (!e) e = bupload(ctx, bt, &d_item_lane, item_lane);
    if (!e) e = bupload(ctx, bt, &d_item_start, item_start);
    if (!e) e = bupload(ctx, bt, &d_lane_off, lane_off);
    if (!e) e = bupload(ctx, bt, &d_lane_len, lane_len);
    if (!e) e = bupload(ctx, bt, &d_item0, bt->h_lane_item0);
    if (!e) e = bupload(ctx, bt, &d_nitem, bt->h_lane_nitem);
    bt->seg = Seg{d_item_lane, d_item_start, d_lane_off, d_lane_len, d_item0, d_nitem, (int64_t)col, bt->nlanes};
    auto dv = [&](double **p, size_t count) { if (!e) e = balloc(ctx, bt, (void **)p, sizeof(double) * count); };
    auto iv = [&](int32_t **p, size_t count) { if (!e) e = balloc(ctx, bt, (void **)p, sizeof(int32_t) * count); };
    dv(&bt->V, (M1 + 1) * col);
    dv(&bt->Z, M1 * col);
    dv(&bt->B, col);
    dv(&bt->X, col);
    dv(&bt->X0, col);
    dv(&bt->R0, col);
    if (!e) e = balloc(ctx, bt, &bt->Y, sizeof(double) * col);
    dv(&bt->part, (size_t)std::max(bt->nitems, 1) * BQ);
    dv(&bt->d1, nl * BD);
    dv(&bt->d2, nl * BD);
    dv(&bt->ycomb, nl * BMAXJ);
    dv(&bt->scale0, nl);
    iv(&bt->live, nl);
    iv(&bt->nv, nl);
    iv(&bt->restart, nl);
    LaneState &st = bt->st;
    iv(&st.active, nl);
    iv(&st.stepped, nl);
    iv(&st.nvecs, nl);
    iv(&st.niter, nl);
    iv(&st.conv, nl);
    iv(&st.nreorth, nl);
    iv(&st.err, nl);
    dv(&st.H, nl * M1 * bt->m);
    dv(&st.cs, nl * M1);
    dv(&st.sn, nl * M1);
    dv(&st.p, nl * M1);
    dv(&st.res, nl);
    dv(&st.r0init, nl);
    dv(&st.hn, nl);
    for (hipEvent_t &v : bt->ev)
        if (!e && hipEventCreateWithFlags(&v, hipEventDisableTiming) != hipSuccess) e = cdg_fail(ctx, CDG_ERR_HIP, "event");
    if (!e && hipHostMalloc((void **)&bt->h_act, sizeof(int32_t) * 3 * nl) != hipSuccess)
        e = cdg_fail(ctx, CDG_ERR_OOM, "batch: pinned flags");
    if (!e) e = bwait(ctx, hipStreamSynchronize(ctx->stream), "allocation");
    const double t_alloc = bsecs(t0) - t_join;
    // the joined factors (one factor launch sequence per block size) and their row-slice storage
    for (BatchGroup &g : bt->groups) {
        const auto tg = bclk::now();
        if (!e) e = cdg_prec_update(ctx, g.M, g.U, nullptr);
        if (!e) e = hipStreamSynchronize(ctx->stream) == hipSuccess ? CDG_OK : cdg_fail(ctx, CDG_ERR_HIP, "factor");
        const double tf = bsecs(tg);
        if (!e) e = rs_prepare(ctx, g.M, g.U);
        if (!e) e = bwait(ctx, hipStreamSynchronize(ctx->stream), "row-slice storage");
        if (btrace())
            std::fprintf(stderr, "[batch] group n=%d: %d systems, %d rows, factor %.3f s, row-slice storage %.3f s\n", g.n,
                         (int)g.members.size(), g.U ? g.U->nrows : 0, tf, bsecs(tg) - tf);
    }
    if (!e) e = bwait(ctx, hipStreamSynchronize(ctx->stream), "factor");
    if (btrace())
        std::fprintf(stderr, "[batch] create: join %.3f s, allocate %.3f s, total %.3f s; %d lanes, %d work items\n",
                     t_join, t_alloc, bsecs(t0), bt->nlanes, bt->nitems);
    if (e) { cdg_batch_destroy(ctx, bt); return e; }
    *out = bt;
    return CDG_OK;
}

// New blocks for the same systems (same patterns, e.g. the next Newton step's
// Jacobians): copied into the joined matrices, which are refactored; nothing
// is allocated.  A = NULL refactors the joined matrices as they are.
extern "C" int cdg_batch_update(cdg_ctx *ctx, cdg_batch *bt, cdg_mat *const *A) {
    if (!ctx || !bt) return CDG_ERR_ARG;
    for (int s = 0; A && s < bt->nsys; ++s)
        if (!A[s] || A[s]->nnzb != bt->sys_nnzb[s] || A[s]->nrows != bt->sys_nrows[s] || A[s]->n != bt->sys_n[s] ||
            !A[s]->part.empty() || A[s]->row_ptr != bt->sys_rp[s] || A[s]->col != bt->sys_col[s] ||
            A[s]->kind != bt->sys_kind[s])
            return cdg_fail(ctx, CDG_ERR_ARG, "batch update: system %d does not match the batch's pattern", s);
    CDG_HIP_TRY(hipDeviceSynchronize());   // the callers' matrices are complete
    const auto t0 = bclk::now();
    int e = CDG_OK;
    for (BatchGroup &g : bt->groups) {
        size_t b0 = 0;
        for (int s : g.members) {
            if (!A) break;
            const size_t bytes = sizeof(double) * (size_t)A[s]->nnzb * A[s]->ld;
            if (!e && hipMemcpyAsync(g.U->blocks + b0 * g.U->ld, A[s]->blocks, bytes, hipMemcpyDeviceToDevice,
                                     ctx->stream) != hipSuccess)
                e = cdg_fail(ctx, CDG_ERR_HIP, "batch update: copying the blocks of system %d", s);
            b0 += A[s]->nnzb;
        }
        if (!e) e = cdg_prec_update(ctx, g.M, g.U, nullptr);
        if (!e) e = rs_prepare(ctx, g.M, g.U);
    }
    if (!e) CDG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (btrace()) std::fprintf(stderr, "[batch] update: copies and factors %.3f s\n", bsecs(t0));
    return e;
}

extern "C" int cdg_batch_destroy(cdg_ctx *ctx, cdg_batch *bt) {
    if (!bt) return CDG_OK;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    batch_drop_graphs(bt);
    for (hipEvent_t &v : bt->ev)
        if (v) (void)hipEventDestroy(v);
    if (bt->h_act) (void)hipHostFree(bt->h_act);
    for (BatchGroup &g : bt->groups) {
        if (g.M) cdg_prec_destroy(ctx, g.M);
        if (g.U) cdg_mat_destroy(ctx, g.U);
    }
    if (ctx)
        for (DevMem &m : bt->owned) dev_free(ctx, m.p, m.bytes);
    delete bt;
    return CDG_OK;
}

extern "C" int cdg_batch_solve(cdg_ctx *ctx, cdg_batch *bt, const double *const *b, double *const *x,
                               const cdg_fgmres_params *prm, cdg_fgmres_stats *stats) {
    if (!ctx || !bt || !b || !x || !prm) return CDG_ERR_ARG;
    if (prm->nkrylov != bt->m) return cdg_fail(ctx, CDG_ERR_ARG, "batch: nkrylov %d, created with %d", prm->nkrylov, bt->m);
    if (prm->inner_fgmres || prm->orthogonalization != CDG_CGS)
        return cdg_fail(ctx, CDG_ERR_UNSUPPORTED, "batch: CGS without inner FGMRES only");
    const auto t_start = bclk::now();
    const int m = bt->m, nl = bt->nlanes, K = bt->K;
    const size_t M1 = (size_t)m + 1, col = (size_t)K * bt->Ltot;
    const double tol = prm->tol, rtol = prm->rtol;
    const int nmax = prm->nmax;
    const bool use_graph = !(std::getenv("CDG_BATCH_GRAPH") && std::atoi(std::getenv("CDG_BATCH_GRAPH")) == 0);
    int e = CDG_OK;
    // right-hand sides in, x and x0 zero
    e = kv_zero(ctx, (int64_t)col, bt->B);
    if (!e) e = kv_zero(ctx, (int64_t)col, bt->X);
    if (!e) e = kv_zero(ctx, (int64_t)col, bt->X0);
    for (int l = 0; l < nl && !e; ++l) {
        const int s = bt->lane_sys[l], r = bt->lane_r[l];
        if (hipMemcpyAsync(bt->B + (size_t)r * bt->Ltot + bt->off[s], b[s] + (size_t)r * bt->len[s],
                           sizeof(double) * bt->len[s], hipMemcpyDefault, ctx->stream) != hipSuccess)
            e = cdg_fail(ctx, CDG_ERR_HIP, "batch: right-hand side of system %d", s);
    }
    const Seg &s = bt->seg;
    const dim3 gi(bt->nitems), gl(nl);
    std::vector<int32_t> h_done(nl, 0), h_live(nl, 0), h_i(nl), h_nrest(nl, 0);
    std::vector<double> h_d(nl * (size_t)BD), h_r0init(nl, 0.0), h_res(nl, std::numeric_limits<double>::max());
    std::vector<int32_t> h_niter(nl, 0), h_nre(nl, 0);
    std::vector<double> h_H(nl * M1 * m), h_p(nl * M1), h_y(nl * (size_t)BMAXJ, 0.0);
    std::vector<int32_t> h_nv(nl, 0), h_conv(nl, 0), h_err(nl, 0), h_restart(nl, 0);
    // lane counters start at 0 (the iteration count spans restarts)
    if (!e) {
        CDG_HIP_TRY(hipMemsetAsync(bt->st.niter, 0, sizeof(int32_t) * nl, ctx->stream));
        CDG_HIP_TRY(hipMemsetAsync(bt->st.nreorth, 0, sizeof(int32_t) * nl, ctx->stream));
        CDG_HIP_TRY(hipMemsetAsync(bt->st.err, 0, sizeof(int32_t) * nl, ctx->stream));
    }
    auto any_running = [&] { for (int l = 0; l < nl; ++l) if (!h_done[l]) return true; return false; };
    while (!e && any_running()) {
        // r0 = b - A x0 for the running lanes; ||r0|| per lane   (type_fgmres.f90:160-163)
        for (BatchGroup &g : bt->groups)
            if (!e) e = group_mv(ctx, bt, g, bt->X0, bt->R0, g.d_row_k);
        if (e) break;
        hipLaunchKernelGGL(k_bsub, dim3((unsigned)((col + 255) / 256)), dim3(256), 0, ctx->stream, (int64_t)col, bt->B,
                           bt->R0);
        for (int l = 0; l < nl; ++l) h_live[l] = !h_done[l];
        CDG_HIP_TRY(hipMemcpyAsync(bt->live, h_live.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_bdots, gi, dim3(BT), 0, ctx->stream, s, bt->V, bt->R0, 0, 1, bt->live, nullptr, bt->part);
        hipLaunchKernelGGL(k_breduce, gl, dim3(64), 0, ctx->stream, s, bt->part, 0, 1, bt->live, nullptr, bt->d1, 0,
                           nullptr);
        CDG_HIP_TRY(hipGetLastError());
        CDG_HIP_TRY(hipMemcpyAsync(h_d.data(), bt->d1, sizeof(double) * h_d.size(), hipMemcpyDeviceToHost, ctx->stream));
        e = ctx_wait(ctx);
        if (e) break;
        std::fill(h_H.begin(), h_H.end(), 0.0);
        std::fill(h_p.begin(), h_p.end(), 0.0);
        std::vector<double> f(nl, 1.0);
        for (int l = 0; l < nl; ++l) {
            h_i[l] = 0;
            if (h_done[l]) continue;
            const double r0 = std::sqrt(h_d[(size_t)l * BD]);
            if (!std::isfinite(r0)) { e = cdg_fail(ctx, CDG_ERR_NAN, "non-finite residual norm (lane %d)", l); break; }
            if (h_niter[l] == 0) h_r0init[l] = r0;
            if (r0 == 0.0) {   // exact x0: x = x0 (the reference divides by zero here)
                h_done[l] = 1;
                h_res[l] = 0.0;
                continue;
            }
            h_i[l] = 1;
            f[l] = r0;
            h_p[(size_t)l * M1] = r0;
        }
        if (e) break;
        // the cycle's lane state: H, rotations, p zeroed, p1 = ||r0||, flags
        CDG_HIP_TRY(hipMemcpyAsync(bt->st.active, h_i.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(bt->st.H, h_H.data(), sizeof(double) * h_H.size(), hipMemcpyHostToDevice, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(bt->st.p, h_p.data(), sizeof(double) * h_p.size(), hipMemcpyHostToDevice, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(bt->st.r0init, h_r0init.data(), sizeof(double) * nl, hipMemcpyHostToDevice,
                                   ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(bt->scale0, f.data(), sizeof(double) * nl, hipMemcpyHostToDevice, ctx->stream));
        CDG_HIP_TRY(hipMemsetAsync(bt->st.cs, 0, sizeof(double) * nl * M1, ctx->stream));
        CDG_HIP_TRY(hipMemsetAsync(bt->st.sn, 0, sizeof(double) * nl * M1, ctx->stream));
        CDG_HIP_TRY(hipMemsetAsync(bt->st.conv, 0, sizeof(int32_t) * nl, ctx->stream));
        CDG_HIP_TRY(hipMemsetAsync(bt->st.nvecs, 0, sizeof(int32_t) * nl, ctx->stream));
        // v1 = r0 / ||r0||
        hipLaunchKernelGGL(k_bdiv, gi, dim3(BT), 0, ctx->stream, s, bt->R0, bt->V, bt->scale0, 1, bt->st.active);
        CDG_HIP_TRY(hipGetLastError());
        // nkrylov Arnoldi steps, each a captured graph (built at the first cycle) or the same
        // launches.  After each step the lanes' active flags go to pinned memory behind an
        // event; the host stays at most two steps ahead of the device and launches no further
        // step once a finished one shows every lane stopped.
        if (use_graph && !bt->execs.empty() &&
            (bt->graph_tol != tol || bt->graph_rtol != rtol || bt->graph_nmax != nmax))
            batch_drop_graphs(bt);   // captured with another stopping test
        if (use_graph && bt->execs.empty()) {
            const auto tc = bclk::now();
            size_t nodes = 0;
            bt->graphs.assign(m, nullptr);
            bt->execs.assign(m, nullptr);
            for (int j = 0; j < m && !e; ++j) {
                CDG_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                const int ce = batch_step(ctx, bt, j, tol, rtol, nmax);
                hipGraph_t gph = nullptr;
                const hipError_t he = hipStreamEndCapture(ctx->stream, &gph);
                if (ce || he != hipSuccess) {
                    if (gph) (void)hipGraphDestroy(gph);
                    e = ce ? ce : cdg_fail(ctx, CDG_ERR_HIP, "graph capture: %s", hipGetErrorString(he));
                    break;
                }
                bt->graphs[j] = gph;
                if (hipGraphInstantiate(&bt->execs[j], gph, nullptr, nullptr, 0) != hipSuccess)
                    e = cdg_fail(ctx, CDG_ERR_HIP, "graph instantiate (step %d)", j);
                size_t nn = 0;
                (void)hipGraphGetNodes(gph, nullptr, &nn);
                nodes += nn;
            }
            if (e) { batch_drop_graphs(bt); break; }
            bt->graph_tol = tol;
            bt->graph_rtol = rtol;
            bt->graph_nmax = nmax;
            if (btrace())
                std::fprintf(stderr, "[batch] graphs: %d steps, %zu nodes, capture + instantiate %.3f s\n", m, nodes,
                             bsecs(tc));
        }
        int launched = 0;
        for (int j = 0; j < m && !e; ++j) {
            if (j >= 2) {
                const int q = (j - 2) % 3;
                CDG_HIP_TRY(hipEventSynchronize(bt->ev[q]));
                bool any = false;
                for (int l = 0; l < nl && !any; ++l) any = bt->h_act[(size_t)q * nl + l] != 0;
                if (!any) break;
            }
            if (use_graph) CDG_HIP_TRY(hipGraphLaunch(bt->execs[j], ctx->stream));
            else e = batch_step(ctx, bt, j, tol, rtol, nmax);
            if (e) break;
            CDG_HIP_TRY(hipMemcpyAsync(bt->h_act + (size_t)(j % 3) * nl, bt->st.active, sizeof(int32_t) * nl,
                                       hipMemcpyDeviceToHost, ctx->stream));
            CDG_HIP_TRY(hipEventRecord(bt->ev[j % 3], ctx->stream));
            ++launched;
        }
        if (e) break;
        // cycle end: y = H^-1 p per lane, x = x0 + Z y; the stopped lanes are done, the others restart
        CDG_HIP_TRY(hipMemcpyAsync(h_nv.data(), bt->st.nvecs, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(h_conv.data(), bt->st.conv, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(h_err.data(), bt->st.err, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(h_niter.data(), bt->st.niter, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(h_res.data(), bt->st.res, sizeof(double) * nl, hipMemcpyDeviceToHost, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(h_H.data(), bt->st.H, sizeof(double) * h_H.size(), hipMemcpyDeviceToHost, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(h_p.data(), bt->st.p, sizeof(double) * h_p.size(), hipMemcpyDeviceToHost, ctx->stream));
        e = ctx_wait(ctx);
        if (e) break;
        if (btrace()) {
            int run = 0, mx = 0;
            for (int l = 0; l < nl; ++l) { run += h_i[l]; mx = std::max(mx, h_nv[l]); }
            std::fprintf(stderr, "[batch] cycle: %d lanes ran, at most %d steps each, %d steps launched; %.3f s since "
                         "the solve began\n", run, mx, launched, bsecs(t_start));
        }
        for (int l = 0; l < nl; ++l)
            if (h_err[l]) { e = cdg_fail(ctx, CDG_ERR_NAN, "non-finite FGMRES residual estimate (lane %d)", l); break; }
        if (e) break;
        for (int l = 0; l < nl; ++l) {
            h_restart[l] = 0;
            if (!h_i[l]) { h_nv[l] = 0; continue; }   // not in this cycle
            const int k = h_nv[l];
            const double *H = h_H.data() + (size_t)l * M1 * m;
            double *y = h_y.data() + (size_t)l * BMAXJ;
            for (int i = k - 1; i >= 0; --i) {
                double acc = h_p[(size_t)l * M1 + i];
                for (int q = i + 1; q < k; ++q) acc -= H[(size_t)q * M1 + i] * y[q];
                y[i] = acc / H[(size_t)i * M1 + i];
            }
            if (h_conv[l] || !(h_res[l] > tol)) {
                h_done[l] = 1;
            } else {
                h_restart[l] = 1;
                h_nrest[l]++;
            }
        }
        CDG_HIP_TRY(hipMemcpyAsync(bt->ycomb, h_y.data(), sizeof(double) * h_y.size(), hipMemcpyHostToDevice, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(bt->nv, h_nv.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(bt->restart, h_restart.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice,
                                   ctx->stream));
        hipLaunchKernelGGL(k_bcombine, gi, dim3(BT), 0, ctx->stream, s, bt->Z, bt->X, bt->X0, bt->ycomb, bt->nv,
                           bt->restart);
        CDG_HIP_TRY(hipGetLastError());
    }
    // ||b - A x|| per lane   (type_fgmres.f90:305,515); x out
    std::vector<double> h_true(nl, 0.0);
    if (!e) {
        for (BatchGroup &g : bt->groups)
            if (!e) e = group_mv(ctx, bt, g, bt->X, bt->R0, g.d_row_k);
    }
    if (!e) {
        hipLaunchKernelGGL(k_bsub, dim3((unsigned)((col + 255) / 256)), dim3(256), 0, ctx->stream, (int64_t)col, bt->B,
                           bt->R0);
        std::vector<int32_t> all(nl, 1);
        CDG_HIP_TRY(hipMemcpyAsync(bt->live, all.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_bdots, gi, dim3(BT), 0, ctx->stream, s, bt->V, bt->R0, 0, 1, bt->live, nullptr, bt->part);
        hipLaunchKernelGGL(k_breduce, gl, dim3(64), 0, ctx->stream, s, bt->part, 0, 1, bt->live, nullptr, bt->d1, 0,
                           nullptr);
        CDG_HIP_TRY(hipGetLastError());
        CDG_HIP_TRY(hipMemcpyAsync(h_d.data(), bt->d1, sizeof(double) * h_d.size(), hipMemcpyDeviceToHost, ctx->stream));
        CDG_HIP_TRY(hipMemcpyAsync(h_nre.data(), bt->st.nreorth, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, ctx->stream));
        for (int l = 0; l < nl && !e; ++l) {
            const int sy = bt->lane_sys[l], r = bt->lane_r[l];
            if (hipMemcpyAsync(x[sy] + (size_t)r * bt->len[sy], bt->X + (size_t)r * bt->Ltot + bt->off[sy],
                               sizeof(double) * bt->len[sy], hipMemcpyDefault, ctx->stream) != hipSuccess)
                e = cdg_fail(ctx, CDG_ERR_HIP, "batch: solution of system %d", sy);
        }
        if (!e) e = ctx_wait(ctx);
        for (int l = 0; l < nl; ++l) h_true[l] = std::sqrt(h_d[(size_t)l * BD]);
    }
    if (stats) {
        const double t = std::chrono::duration<double>(bclk::now() - t_start).count();
        for (int l = 0; l < nl; ++l) {
            cdg_fgmres_stats &o = stats[l];
            std::memset(&o, 0, sizeof(o));
            o.niter = h_niter[l];
            o.residual_norm = h_true[l];
            o.nrestart = h_nrest[l];
            o.nreorth = h_nre[l];
            o.t_total = t;
        }
    }
    return e;
}

// Lanes in stats order: system s's right-hand sides r = 0 .. nrhs[s]-1, systems in order.
extern "C" int cdg_batch_info(const cdg_batch *bt, int32_t *nlanes, int32_t *ngroups, int64_t *values) {
    if (!bt) return CDG_ERR_ARG;
    if (nlanes) *nlanes = bt->nlanes;
    if (ngroups) *ngroups = (int32_t)bt->groups.size();
    if (values) *values = bt->Ltot;
    return CDG_OK;
}
