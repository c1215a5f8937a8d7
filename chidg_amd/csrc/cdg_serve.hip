// Request loop for parameter-sweep solves on one GPU (no reference
// counterpart: ChiDG solves one system per MPI rank).
//
// A server owns `nslots` worker threads, each with its own context (HIP
// stream, reduction scratch, cached device memory).  Requests queue in
// arrival order; a free worker admits the request at the head of the queue
// as soon as its device-memory footprint fits the budget next to the
// requests already running, so requests that arrive while others run start
// as slots and memory free up, and their work overlaps on the device (one
// stream per slot).  A request's Jacobian, ILU(0) factor and Krylov basis
// stay on the device for the whole solve; its result depends only on the
// request (every kernel sums in a fixed order), not on what runs beside it.
#include "cdg_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

// Uniform hex box in ChiDG row order: element i fastest; per row the interior
// neighbours in face order XI_MIN, XI_MAX, ETA_MIN, ETA_MAX, ZETA_MIN, ZETA_MAX,
// then the diagonal, then periodic wrap neighbours as coupling-only blocks
// (a donor already present is not added twice).  Same order as
// chidg_amd/pattern.py hex_box (type_domain_matrix.f90:134,253-374).  face
// (optional): 1..6 for the face an entry came from (a coupling entry: the
// first face that found its donor), 7 for the diagonal.
void cdg_hex_box(int nx, int ny, int nz, int periodic, std::vector<int32_t> &row_ptr, std::vector<int32_t> &col,
                 std::vector<int32_t> &kind, std::vector<int32_t> *face) {
    const int dims[3] = {nx, ny, nz};
    const int stride[3] = {1, nx, nx * ny};
    const int nrows = nx * ny * nz;
    row_ptr.assign(nrows + 1, 0);
    col.clear();
    kind.clear();
    if (face) face->clear();
    for (int e = 0; e < nrows; ++e) {
        const int idx[3] = {e % nx, (e / nx) % ny, e / (nx * ny)};
        std::vector<std::pair<int, int>> chi;   // (donor, face)
        for (int d = 0; d < 3; ++d)
            for (int side = -1; side <= 1; side += 2) {
                const int fid = 1 + 2 * d + (side > 0);
                const int q = idx[d] + side;
                if (q >= 0 && q < dims[d]) {
                    col.push_back(e + side * stride[d]);
                    kind.push_back(CDG_KIND_LOCAL);
                    if (face) face->push_back(fid);
                } else if (periodic & (1 << d)) {
                    const int wrap = (q + dims[d]) % dims[d];
                    const int donor = e + (wrap - idx[d]) * stride[d];
                    bool seen = false;
                    for (auto &c : chi) seen = seen || c.first == donor;
                    if (!seen) chi.push_back({donor, fid});
                }
            }
        col.push_back(e);
        kind.push_back(CDG_KIND_LOCAL);
        if (face) face->push_back(7);
        for (auto &c : chi) {
            col.push_back(c.first);
            kind.push_back(CDG_KIND_COUPLING);
            if (face) face->push_back(c.second);
        }
        row_ptr[e + 1] = (int32_t)col.size();
    }
}

namespace {

using clk = std::chrono::steady_clock;

double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

int block_n(const cdg_request &r) { return (r.order + 1) * (r.order + 1) * (r.order + 1) * 5; }

// A caller's matrix registered with the server: its ILU(0) factor lives on
// the home worker's context from the first request on and is reused while
// the matrix stamp is unchanged (solver_controller_t,
// type_solver_controller.f90:89-124).
struct System {
    const cdg_mat *A = nullptr;        // the caller's (not owned)
    int32_t precision = CDG_FP64;
    int home = 0;                      // worker slot that runs its requests
    cdg_prec *M = nullptr;             // on the home worker's context
    int64_t resident = 0;              // bytes reserved while registered (the factor)
};

enum JobKind { JOB_SYNTHETIC = 0, JOB_SYSTEM = 1, JOB_RELEASE = 2 };

struct Job {
    uint64_t ticket;
    int kind = JOB_SYNTHETIC;
    cdg_request req;
    uint64_t system = 0;               // JOB_SYSTEM / JOB_RELEASE
    int home = -1;                     // worker slot that must run it (-1: any)
    const double *b = nullptr;         // JOB_SYSTEM: nrhs right-hand sides
    int32_t nrhs = 1;
    double *x_out;
    int64_t footprint;
    clk::time_point t_submit;
};

struct Slot {
    cdg_result res{};
    bool done = false;
};

}  // namespace

struct cdg_server {
    int device = 0;
    int64_t budget = 0;
    cdg_fgmres_params params{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> queue;
    std::unordered_map<uint64_t, Slot> results;
    uint64_t next_ticket = 1;
    int64_t in_use = 0;
    int active = 0, peak_active = 0;
    bool stop = false;
    int nslots = 0;
    std::vector<std::thread> workers;
    std::string err;
    std::unordered_map<uint64_t, System> systems;
    uint64_t next_system = 1;
};

// Device bytes of an ILU(0) preconditioner beyond the matrix: the factor (L,
// Dinv), the FP32/BF16 copies of the upper blocks, the BF16 variant's FP32
// work, the block inverse's working slots, the level schedule's partial sums,
// and the stream schedule's row-slice copies (L, W = Dinv U, Dinv) with its
// tagged exchange vectors (nup = nlo: the kind-0 pattern is symmetric).
static int64_t factor_extra_bytes(int64_t nrows, int64_t nlo, int64_t n, int precision) {
    const int64_t ld = slot_elems((int)n), len = nrows * n;
    const int64_t es = precision == CDG_FP64 ? 8 : (precision == CDG_FP32 ? 4 : 2);
    int64_t f = es * (nlo + nrows) * ld;                         // L, Dinv
    if (precision != CDG_FP64) f += es * nlo * ld;               // U copy
    if (precision == CDG_BF16) f += 4 * (nlo + nrows) * ld;      // FP32 factor work
    if (n > 192) f += 8 * nrows * ld;                            // inverse working slots (bound)
    f += 3 * 8 * nrows * 8 * n;                                  // apply partial sums (<= 8 slices)
    f += es * (2 * nlo + nrows) * (ld + 64 * 16);                // stream schedule slots (slice padding bound)
    f += 2 * 8 * len * (precision == CDG_FP64 ? 2 : 1);          // tagged exchange vectors
    return f;
}

extern "C" int cdg_server_footprint(const cdg_request *req, const cdg_fgmres_params *params, int64_t *bytes) {
    if (!req || !params || !bytes) return CDG_ERR_ARG;
    if (req->nx <= 0 || req->ny <= 0 || req->nz <= 0 || req->order < 0 || req->order > 3) return CDG_ERR_ARG;
    std::vector<int32_t> rp, col, kind;
    cdg_hex_box(req->nx, req->ny, req->nz, req->periodic, rp, col, kind);
    const int64_t nrows = (int64_t)rp.size() - 1, nnzb = (int64_t)col.size();
    int64_t nloc = 0;
    for (int32_t k : kind) nloc += (k == CDG_KIND_LOCAL);
    const int64_t n = block_n(*req), ld = slot_elems((int)n), len = nrows * n;
    const int64_t nlo = (nloc - nrows) / 2;
    int64_t f = 8 * nnzb * ld;                                   // A
    f += factor_extra_bytes(nrows, nlo, n, req->precision);
    const int64_t m = std::max(1, params->nkrylov);
    const int64_t k = std::max(1, req->nrhs);
    f += 8 * len * k * (2 * (m + 1) + 6);                        // V, Z, b, x, work vectors per right-hand side
    if (params->inner_fgmres) f += 8 * len * (2 * (std::max(1, params->inner_nkrylov) + 1) + 4);
    *bytes = f + f / 20 + (64ll << 20);
    return CDG_OK;
}

namespace {

// One request, start to finish, on the worker's context.
int run_job(cdg_ctx *ctx, const cdg_fgmres_params &params, Job &job, cdg_result &res) {
    const cdg_request &r = job.req;
    std::vector<int32_t> rp, col, kind;
    cdg_hex_box(r.nx, r.ny, r.nz, r.periodic, rp, col, kind);
    const int nrows = (int)rp.size() - 1, n = block_n(r);
    const int64_t len = (int64_t)nrows * n;
    cdg_mat *A = nullptr;
    cdg_prec *M = nullptr;
    double *b = nullptr, *x = nullptr;
    const int nrhs = std::max(1, r.nrhs);
    int e = cdg_mat_create(ctx, nrows, n, rp.data(), col.data(), kind.data(), &A);
    if (!e) e = cdg_mat_fill_synthetic(ctx, A, r.seed_A, r.s_diag, r.s_off);
    if (!e) e = dev_alloc(ctx, (void **)&b, sizeof(double) * len * nrhs);
    if (!e) e = dev_alloc(ctx, (void **)&x, sizeof(double) * len * nrhs);
    // right-hand sides k = 0.. (seeds seed_b + k) share the matrix and the
    // factor (updated once); several of them are solved in lockstep, each
    // block read once per step for all of them
    for (int k = 0; k < nrhs && !e; ++k)
        e = launch_vec_synthetic(ctx, len, r.seed_b + (uint64_t)k, 1.0, b + (size_t)k * len);
    if (!e) e = cdg_prec_create(ctx, A, CDG_PREC_ILU0, r.precision, &M);
    cdg_fgmres_params prm = params;
    prm.update_preconditioner = 1;
    res.niter = 0;
    res.residual_norm = 0.0;
    std::vector<cdg_fgmres_stats> st(nrhs);
    if (!e) e = cdg_fgmres_solve_multi(ctx, A, M, nrhs, b, x, &prm, st.data());
    for (int k = 0; k < nrhs && !e; ++k) {
        double bn = 0.0;
        e = ctx_norm2(ctx, len, b + (size_t)k * len, &bn);
        res.niter += st[k].niter;
        res.residual_norm = std::max(res.residual_norm, bn > 0 ? st[k].residual_norm / bn : st[k].residual_norm);
    }
    if (!e) res.t_factor = st[0].t_factor;
    if (!e && job.x_out) {
        if (hipMemcpyAsync(job.x_out, x, sizeof(double) * len, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            e = cdg_fail(ctx, CDG_ERR_HIP, "copying the solution out failed");
    }
    if (M) cdg_prec_destroy(ctx, M);
    if (A) cdg_mat_destroy(ctx, A);
    dev_free(ctx, b, 0);
    dev_free(ctx, x, 0);
    return e;
}

bool on_device(const void *p) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice;
}

// ||b|| of a host or device vector
int vec_norm(cdg_ctx *ctx, const double *b, int64_t len, double *out) {
    if (on_device(b)) return ctx_norm2(ctx, len, b, out);
    double t = 0.0;
    for (int64_t k = 0; k < len; ++k) t += b[k] * b[k];
    *out = std::sqrt(t);
    return CDG_OK;
}

// A request on a registered system, on its home worker: the factor is
// created at the first request and refactored only when the matrix stamp
// changed since (update_preconditioner = 0: cdg_fgmres_solve compares the
// stamps); the right-hand sides share it.
int run_system_job(cdg_server *s, cdg_ctx *ctx, Job &job, cdg_result &res) {
    System *sys;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->systems.find(job.system);
        if (it == s->systems.end()) return cdg_fail(ctx, CDG_ERR_ARG, "system %llu is not registered",
                                                    (unsigned long long)job.system);
        sys = &it->second;   // only its home worker (this thread) touches M
    }
    const cdg_mat *A = sys->A;
    int e = CDG_OK;
    if (!sys->M) e = cdg_prec_create(ctx, A, CDG_PREC_ILU0, sys->precision, &sys->M);
    const int64_t len = (int64_t)A->nrows * A->n;
    double *xs = nullptr;
    if (!e && !job.x_out)
        e = dev_alloc(ctx, (void **)&xs, sizeof(double) * (size_t)std::max<int64_t>(len, 1) * std::max(1, job.nrhs));
    cdg_fgmres_params prm = s->params;
    prm.update_preconditioner = 0;
    res.niter = 0;
    res.residual_norm = 0.0;
    res.factored = 0;
    const int nrhs = std::max(1, job.nrhs);
    std::vector<cdg_fgmres_stats> st(nrhs);
    if (!e) e = cdg_fgmres_solve_multi(ctx, A, sys->M, nrhs, job.b, job.x_out ? job.x_out : xs, &prm, st.data());
    for (int k = 0; k < nrhs && !e; ++k) {
        double bn = 0.0;
        e = vec_norm(ctx, job.b + (size_t)k * len, len, &bn);
        res.niter += st[k].niter;
        res.residual_norm = std::max(res.residual_norm, bn > 0 ? st[k].residual_norm / bn : st[k].residual_norm);
    }
    if (!e) {
        res.t_factor = st[0].t_factor;
        res.factored = st[0].factored;
    }
    if (xs) dev_free(ctx, xs, 0);
    return e;
}

// Unregistering: the factor goes with the home worker's context.
int release_system(cdg_server *s, cdg_ctx *ctx, Job &job) {
    cdg_prec *M = nullptr;
    int64_t resident = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->systems.find(job.system);
        if (it == s->systems.end()) return CDG_OK;
        M = it->second.M;
        resident = it->second.resident;
        s->systems.erase(it);
        s->in_use -= resident;
    }
    if (M) cdg_prec_destroy(ctx, M);
    return CDG_OK;
}

void worker(cdg_server *s, int slot) {
    cdg_ctx *ctx = nullptr;
    const int ce = cdg_ctx_create(s->device, &ctx);
    for (;;) {
        Job job;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // the first queued job this worker may run (any worker, or this
            // one as a registered system's home), admitted when it fits
            std::deque<Job>::iterator pick;
            s->cv_work.wait(lk, [&] {
                pick = s->queue.end();
                for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
                    if (it->home < 0 || it->home == slot) { pick = it; break; }
                if (s->stop && pick == s->queue.end()) return true;
                if (pick == s->queue.end()) return false;
                return s->in_use + pick->footprint <= s->budget || s->active == 0;
            });
            if (pick == s->queue.end()) break;
            job = *pick;
            s->queue.erase(pick);
            s->in_use += job.footprint;
            s->active += 1;
            s->peak_active = std::max(s->peak_active, s->active);
        }
        cdg_result res{};
        res.slot = slot;
        const clk::time_point t_start = clk::now();
        res.t_queue = secs(job.t_submit, t_start);
        int e;
        if (ce) e = ce;
        else if (job.footprint > s->budget) e = CDG_ERR_OOM;
        else if (job.kind == JOB_SYSTEM) e = run_system_job(s, ctx, job, res);
        else if (job.kind == JOB_RELEASE) e = release_system(s, ctx, job);
        else e = run_job(ctx, s->params, job, res);
        const clk::time_point t_end = clk::now();
        res.status = e;
        res.t_run = secs(t_start, t_end);
        res.t_latency = secs(job.t_submit, t_end);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->in_use -= job.footprint;
            s->active -= 1;
            Slot &sl = s->results[job.ticket];
            sl.res = res;
            sl.done = true;
            if (e && ctx) s->err = ctx->err;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    {   // factors of systems still registered at shutdown live on this context
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto &kv : s->systems)
            if (kv.second.home == slot && kv.second.M) {
                cdg_prec_destroy(ctx, kv.second.M);
                kv.second.M = nullptr;
            }
    }
    if (ctx) cdg_ctx_destroy(ctx);
}

}  // namespace

extern "C" int cdg_server_create(int32_t device, int32_t nslots, int64_t hbm_budget, const cdg_fgmres_params *params,
                                 cdg_server **out) {
    if (!out || nslots < 1 || nslots > 64) return CDG_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { (void)hipGetLastError(); return CDG_ERR_HIP; }
    if (device < 0 || device >= ndev) return CDG_ERR_ARG;
    cdg_server *s = new cdg_server();
    s->device = device;
    if (params) s->params = *params;
    else cdg_fgmres_default_params(&s->params);
    if (hbm_budget <= 0) {
        size_t fr = 0, tot = 0;
        (void)hipSetDevice(device);
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete s; return CDG_ERR_HIP; }
        hbm_budget = (int64_t)(0.8 * (double)fr);
    }
    s->budget = hbm_budget;
    s->nslots = nslots;
    for (int i = 0; i < nslots; ++i) s->workers.emplace_back(worker, s, i);
    *out = s;
    return CDG_OK;
}

extern "C" int cdg_server_submit(cdg_server *s, const cdg_request *req, double *x_out, uint64_t *ticket) {
    if (!s || !req || !ticket) return CDG_ERR_ARG;
    Job job;
    job.req = *req;
    job.x_out = x_out;
    job.t_submit = clk::now();
    int e = cdg_server_footprint(req, &s->params, &job.footprint);
    if (e) return e;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return CDG_ERR_STATE;
        job.ticket = s->next_ticket++;
        s->results[job.ticket] = Slot{};
        s->queue.push_back(job);
        *ticket = job.ticket;
    }
    s->cv_work.notify_all();
    return CDG_OK;
}

extern "C" int cdg_server_register(cdg_server *s, const cdg_mat *A, int32_t precision, uint64_t *system) {
    if (!s || !A || !system) return CDG_ERR_ARG;
    if (precision < CDG_FP64 || precision > CDG_BF16) return CDG_ERR_ARG;
    if (A->struct_err != CDG_OK) return A->struct_err;
    System sys;
    sys.A = A;
    sys.precision = precision;
    sys.resident = factor_extra_bytes(A->nrows, (int64_t)A->lo_blk.size(), A->n, precision);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return CDG_ERR_STATE;
    if (s->in_use + sys.resident > s->budget) return CDG_ERR_OOM;
    const uint64_t id = s->next_system++;
    sys.home = (int)((id - 1) % (uint64_t)s->nslots);
    s->in_use += sys.resident;
    s->systems[id] = sys;
    *system = id;
    return CDG_OK;
}

extern "C" int cdg_server_submit_system(cdg_server *s, uint64_t system, const double *b, int32_t nrhs, double *x_out,
                                        uint64_t *ticket) {
    if (!s || !b || !ticket || nrhs < 1) return CDG_ERR_ARG;
    Job job;
    job.kind = JOB_SYSTEM;
    job.system = system;
    job.b = b;
    job.nrhs = nrhs;
    job.x_out = x_out;
    job.t_submit = clk::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return CDG_ERR_STATE;
        auto it = s->systems.find(system);
        if (it == s->systems.end()) return CDG_ERR_ARG;
        job.home = it->second.home;
        const cdg_mat *A = it->second.A;
        const int64_t len = (int64_t)A->nrows * A->n, m = std::max(1, s->params.nkrylov);
        job.footprint = 8 * len * (2 * (m + 1) + 8);            // Krylov basis, staging, work vectors
        if (s->params.inner_fgmres) job.footprint += 8 * len * (2 * (std::max(1, s->params.inner_nkrylov) + 1) + 4);
        job.ticket = s->next_ticket++;
        s->results[job.ticket] = Slot{};
        s->queue.push_back(job);
        *ticket = job.ticket;
    }
    s->cv_work.notify_all();
    return CDG_OK;
}

extern "C" int cdg_server_unregister(cdg_server *s, uint64_t system) {
    if (!s) return CDG_ERR_ARG;
    Job job;
    job.kind = JOB_RELEASE;
    job.system = system;
    job.x_out = nullptr;
    job.footprint = 0;
    job.t_submit = clk::now();
    uint64_t ticket;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->systems.find(system);
        if (it == s->systems.end()) return CDG_ERR_ARG;
        job.home = it->second.home;
        job.ticket = ticket = s->next_ticket++;
        s->results[ticket] = Slot{};
        s->queue.push_back(job);   // behind the system's queued requests
    }
    s->cv_work.notify_all();
    cdg_result res{};
    return cdg_server_wait(s, ticket, &res);
}

extern "C" int cdg_server_wait(cdg_server *s, uint64_t ticket, cdg_result *res) {
    if (!s || !res) return CDG_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return CDG_ERR_ARG;
    s->cv_done.wait(lk, [&] { return s->results[ticket].done; });
    *res = s->results[ticket].res;
    s->results.erase(ticket);
    return CDG_OK;
}

extern "C" int cdg_server_info(cdg_server *s, int32_t *queued, int32_t *active, int32_t *peak_active,
                               int64_t *in_use, int64_t *budget) {
    if (!s) return CDG_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued) *queued = (int32_t)s->queue.size();
    if (active) *active = s->active;
    if (peak_active) *peak_active = s->peak_active;
    if (in_use) *in_use = s->in_use;
    if (budget) *budget = s->budget;
    return CDG_OK;
}

extern "C" const char *cdg_server_last_error(cdg_server *s) { return s ? s->err.c_str() : "null server"; }

extern "C" int cdg_server_destroy(cdg_server *s) {
    if (!s) return CDG_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (std::thread &t : s->workers) t.join();
    delete s;
    return CDG_OK;
}
