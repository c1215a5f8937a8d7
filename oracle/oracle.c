/*
 * oracle.c -- CPU FP64 restatement of ChiDG's linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY (checker + CPU baseline).  Never part of the
 * product path.  Written from the reference's behaviour; each function cites
 * the reference lines it restates.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_g_threads = 1; /* shared with synth.c */

void orc_set_threads(int nthreads) { orc_g_threads = nthreads > 0 ? nthreads : 1; }
int orc_get_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* Pattern analysis: diagonal, lower, upper, transpose (kind-0 blocks only)   */
/* type_domain_matrix.f90:295-302 (lower iff parent < row), :592-627 (transpose) */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t *diag;       /* nrows: block index of diagonal                   */
    int32_t *lo_ptr;     /* nrows+1                                          */
    int32_t *lo_blk;     /* lower kind-0 blocks in row order                  */
    int32_t *lo_tr;      /* transpose block (j,i) of each lower block (i,j)   */
    int32_t *up_ptr;
    int32_t *up_blk;
    int32_t *level;      /* forward level of each row                         */
    int32_t *blevel;     /* backward level of each row                        */
    int32_t nlevels, nblevels;
    int32_t *lev_ptr, *lev_rows;    /* rows grouped by forward level          */
    int32_t *blev_ptr, *blev_rows;  /* rows grouped by backward level         */
} orc_struct;

static void struct_free(orc_struct *s)
{
    free(s->diag); free(s->lo_ptr); free(s->lo_blk); free(s->lo_tr);
    free(s->up_ptr); free(s->up_blk); free(s->level); free(s->blevel);
    free(s->lev_ptr); free(s->lev_rows); free(s->blev_ptr); free(s->blev_rows);
    memset(s, 0, sizeof(*s));
}

static int group_levels(int nrows, const int32_t *lev, int nlev, int32_t **ptr_out, int32_t **rows_out)
{
    int32_t *ptr = (int32_t *)calloc((size_t)nlev + 1, sizeof(int32_t));
    int32_t *rows = (int32_t *)malloc((size_t)(nrows > 0 ? nrows : 1) * sizeof(int32_t));
    int32_t *fill = (int32_t *)calloc((size_t)nlev + 1, sizeof(int32_t));
    if (!ptr || !rows || !fill) { free(ptr); free(rows); free(fill); return ORC_ERR_ALLOC; }
    for (int r = 0; r < nrows; ++r) ptr[lev[r] + 1]++;
    for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
    for (int r = 0; r < nrows; ++r) rows[ptr[lev[r]] + fill[lev[r]]++] = r;
    free(fill);
    *ptr_out = ptr; *rows_out = rows;
    return ORC_OK;
}

static int struct_build(const orc_pattern *p, orc_struct *s, int need_transpose)
{
    memset(s, 0, sizeof(*s));
    const int nr = p->nrows;
    s->diag = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nr > 0 ? nr : 1));
    s->lo_ptr = (int32_t *)calloc((size_t)nr + 1, sizeof(int32_t));
    s->up_ptr = (int32_t *)calloc((size_t)nr + 1, sizeof(int32_t));
    s->level = (int32_t *)calloc((size_t)(nr > 0 ? nr : 1), sizeof(int32_t));
    s->blevel = (int32_t *)calloc((size_t)(nr > 0 ? nr : 1), sizeof(int32_t));
    if (!s->diag || !s->lo_ptr || !s->up_ptr || !s->level || !s->blevel) { struct_free(s); return ORC_ERR_ALLOC; }
    for (int i = 0; i < nr; ++i) {
        s->diag[i] = -1;
        for (int b = p->row_ptr[i]; b < p->row_ptr[i + 1]; ++b) {
            if (p->kind[b] != 0) continue;
            int j = p->col[b];
            if (j == i) { if (s->diag[i] < 0) s->diag[i] = b; }
            else if (p->part && p->part[j] != p->part[i]) continue;   /* parent on another rank */
            else if (j < i) s->lo_ptr[i + 1]++;
            else s->up_ptr[i + 1]++;
        }
        if (s->diag[i] < 0) { struct_free(s); return ORC_ERR_ARG; }
    }
    for (int i = 0; i < nr; ++i) { s->lo_ptr[i + 1] += s->lo_ptr[i]; s->up_ptr[i + 1] += s->up_ptr[i]; }
    int nlo = s->lo_ptr[nr], nup = s->up_ptr[nr];
    s->lo_blk = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nlo > 0 ? nlo : 1));
    s->lo_tr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nlo > 0 ? nlo : 1));
    s->up_blk = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nup > 0 ? nup : 1));
    if (!s->lo_blk || !s->lo_tr || !s->up_blk) { struct_free(s); return ORC_ERR_ALLOC; }
    for (int i = 0; i < nr; ++i) {
        int il = s->lo_ptr[i], iu = s->up_ptr[i];
        for (int b = p->row_ptr[i]; b < p->row_ptr[i + 1]; ++b) {
            if (p->kind[b] != 0) continue;
            int j = p->col[b];
            if (p->part && p->part[j] != p->part[i]) continue;
            if (j < i) s->lo_blk[il++] = b;
            else if (j > i) s->up_blk[iu++] = b;
        }
    }
    /* transpose position: first kind-0 block of row j whose parent is i
     * (type_domain_matrix.f90:605-623) */
    for (int i = 0; i < nr; ++i) {
        for (int q = s->lo_ptr[i]; q < s->lo_ptr[i + 1]; ++q) {
            int j = p->col[s->lo_blk[q]];
            int found = -1;
            for (int b = p->row_ptr[j]; b < p->row_ptr[j + 1]; ++b)
                if (p->kind[b] == 0 && p->col[b] == i) { found = b; break; }
            if (found < 0 && need_transpose) { struct_free(s); return ORC_ERR_NO_TRANSPOSE; }
            s->lo_tr[q] = found;
        }
    }
    /* level sets: rows of one level have no dependency on each other */
    int nl = 0;
    for (int i = 0; i < nr; ++i) {
        int l = 0;
        for (int q = s->lo_ptr[i]; q < s->lo_ptr[i + 1]; ++q) {
            int j = p->col[s->lo_blk[q]];
            if (s->level[j] + 1 > l) l = s->level[j] + 1;
        }
        s->level[i] = l;
        if (l + 1 > nl) nl = l + 1;
    }
    int nb = 0;
    for (int i = nr - 1; i >= 0; --i) {
        int l = 0;
        for (int q = s->up_ptr[i]; q < s->up_ptr[i + 1]; ++q) {
            int k = p->col[s->up_blk[q]];
            if (s->blevel[k] + 1 > l) l = s->blevel[k] + 1;
        }
        s->blevel[i] = l;
        if (l + 1 > nb) nb = l + 1;
    }
    s->nlevels = nl; s->nblevels = nb;
    if (group_levels(nr, s->level, nl, &s->lev_ptr, &s->lev_rows) ||
        group_levels(nr, s->blevel, nb, &s->blev_ptr, &s->blev_rows)) { struct_free(s); return ORC_ERR_ALLOC; }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* Dense block kernels (column-major n x n)                                  */
/* ------------------------------------------------------------------------- */

/* tmp = A x ; y += tmp     -- resvec = resvec + matmul(Amat,xvec)
 * (operator_chidg_mv.f90:224-230) */
static void blk_gemv_acc(int n, const double *A, const double *x, double *y, double *tmp)
{
    for (int r = 0; r < n; ++r) tmp[r] = 0.0;
    for (int c = 0; c < n; ++c) {
        const double xc = x[c];
        const double *Ac = A + (size_t)c * n;
        for (int r = 0; r < n; ++r) tmp[r] += Ac[r] * xc;
    }
    for (int r = 0; r < n; ++r) y[r] += tmp[r];
}

/* y = y - A x  (temp = z_i - matmul(L, z_j), precon_ILU0.f90:352) */
static void blk_gemv_sub(int n, const double *A, const double *x, double *y, double *tmp)
{
    for (int r = 0; r < n; ++r) tmp[r] = 0.0;
    for (int c = 0; c < n; ++c) {
        const double xc = x[c];
        const double *Ac = A + (size_t)c * n;
        for (int r = 0; r < n; ++r) tmp[r] += Ac[r] * xc;
    }
    for (int r = 0; r < n; ++r) y[r] -= tmp[r];
}

/* C = A B (all column-major n x n) -- matmul(A,B) */
static void blk_gemm(int n, const double *A, const double *B, double *C)
{
    for (int j = 0; j < n; ++j) {
        double *Cj = C + (size_t)j * n;
        for (int r = 0; r < n; ++r) Cj[r] = 0.0;
        for (int k = 0; k < n; ++k) {
            const double bkj = B[(size_t)j * n + k];
            const double *Ak = A + (size_t)k * n;
            for (int r = 0; r < n; ++r) Cj[r] += Ak[r] * bkj;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* mod_inv::inv  (mod_inv.f90:99-150) via LAPACK's published unblocked        */
/* algorithms: DGETF2 (LU, partial pivoting, first max on ties), DTRTI2       */
/* (upper-triangular inverse), DGETRI (inv(A) from inv(U) and L, then column  */
/* interchanges).                                                              */
/* ------------------------------------------------------------------------- */
int orc_inv(int n, const double *A, double *Ainv)
{
    if (n <= 0) return ORC_ERR_ARG;
    double *a = Ainv;
    if (a != A) memcpy(a, A, sizeof(double) * (size_t)n * n);
    int *ipiv = (int *)malloc(sizeof(int) * (size_t)n);
    double *work = (double *)malloc(sizeof(double) * (size_t)n);
    if (!ipiv || !work) { free(ipiv); free(work); return ORC_ERR_ALLOC; }
#define AA(i, j) a[(size_t)(j) * n + (i)]
    int info = 0;
    /* DGETF2 */
    for (int j = 0; j < n; ++j) {
        int jp = j;
        double amax = fabs(AA(j, j));
        for (int i = j + 1; i < n; ++i)
            if (fabs(AA(i, j)) > amax) { amax = fabs(AA(i, j)); jp = i; }
        ipiv[j] = jp;
        if (AA(jp, j) != 0.0) {
            if (jp != j)
                for (int c = 0; c < n; ++c) { double t = AA(j, c); AA(j, c) = AA(jp, c); AA(jp, c) = t; }
            const double r = 1.0 / AA(j, j);
            for (int i = j + 1; i < n; ++i) AA(i, j) *= r;
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; ++c) {
            const double t = AA(j, c);
            if (t != 0.0)
                for (int i = j + 1; i < n; ++i) AA(i, c) -= AA(i, j) * t;
        }
    }
    if (info != 0 || !isfinite(AA(n - 1, n - 1))) { free(ipiv); free(work); return ORC_ERR_SINGULAR; }
    /* DTRTI2: inv(U) in place, column by column */
    for (int j = 0; j < n; ++j) {
        AA(j, j) = 1.0 / AA(j, j);
        const double ajj = -AA(j, j);
        /* x(0:j) = triu(U(0:j,0:j)) * x   (DTRMV upper, no-trans, non-unit) */
        for (int jj = 0; jj < j; ++jj) {
            const double t = AA(jj, j);
            if (t != 0.0) {
                for (int i = 0; i < jj; ++i) AA(i, j) += t * AA(i, jj);
                AA(jj, j) = t * AA(jj, jj);
            }
        }
        for (int i = 0; i < j; ++i) AA(i, j) *= ajj;
    }
    /* DGETRI unblocked: solve inv(A)*L = inv(U) */
    for (int j = n - 1; j >= 0; --j) {
        for (int i = j + 1; i < n; ++i) { work[i] = AA(i, j); AA(i, j) = 0.0; }
        for (int c = j + 1; c < n; ++c) {
            const double t = work[c];
            for (int i = 0; i < n; ++i) AA(i, j) -= AA(i, c) * t;
        }
    }
    for (int j = n - 2; j >= 0; --j) {
        const int jp = ipiv[j];
        if (jp != j)
            for (int i = 0; i < n; ++i) { double t = AA(i, j); AA(i, j) = AA(i, jp); AA(i, jp) = t; }
    }
#undef AA
    free(ipiv); free(work);
    for (size_t k = 0; k < (size_t)n * n; ++k)
        if (!isfinite(a[k])) return ORC_ERR_SINGULAR;
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* Matvec (operator_chidg_mv.f90:117-646, single rank: every block local)     */
/* ------------------------------------------------------------------------- */
int orc_mv(const orc_pattern *p, const double *blocks, const double *x, double *y)
{
    const int n = p->n;
    const size_t nn = (size_t)n * n;
    int err = ORC_OK;
#pragma omp parallel num_threads(orc_g_threads)
    {
        double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
        if (!tmp) {
#pragma omp atomic write
            err = ORC_ERR_ALLOC;
        } else {
#pragma omp for schedule(static)
            for (int i = 0; i < p->nrows; ++i) {
                double *yi = y + (size_t)i * n;
                for (int r = 0; r < n; ++r) yi[r] = 0.0;
                /* lblks first, then coupling-only blocks: pattern order */
                for (int b = p->row_ptr[i]; b < p->row_ptr[i + 1]; ++b)
                    blk_gemv_acc(n, blocks + (size_t)b * nn, x + (size_t)p->col[b] * n, yi, tmp);
            }
            free(tmp);
        }
    }
    return err;
}

/* ------------------------------------------------------------------------- */
/* ILU0 update (precon_ILU0.f90:111-265, native, A not transposed)           */
/* ------------------------------------------------------------------------- */
int orc_ilu0_factor(const orc_pattern *p, const double *blocks, double *L, double *Dinv, int *bad_row)
{
    const int n = p->n;
    const size_t nn = (size_t)n * n;
    orc_struct s;
    int err = struct_build(p, &s, 1);
    if (err) return err;
    /* LD.diag(e) = A.diag(e)  (:155-159) */
    for (int i = 0; i < p->nrows; ++i)
        memcpy(Dinv + (size_t)i * nn, blocks + (size_t)s.diag[i] * nn, sizeof(double) * nn);
    int bad = -1;
    /* rows of one level are independent; natural order inside a level */
    for (int l = 0; l < s.nlevels && bad < 0; ++l) {
#pragma omp parallel num_threads(orc_g_threads)
        {
            double *tmp = (double *)malloc(sizeof(double) * nn);
            double *inv = (double *)malloc(sizeof(double) * nn);
#pragma omp for schedule(dynamic, 1)
            for (int q = s.lev_ptr[l]; q < s.lev_ptr[l + 1]; ++q) {
                const int i = s.lev_rows[q];
                double *Di = Dinv + (size_t)i * nn;
                if (!tmp || !inv) {
#pragma omp critical
                    { if (bad < 0) { bad = i; err = ORC_ERR_ALLOC; } }
                    continue;
                }
                for (int t = s.lo_ptr[i]; t < s.lo_ptr[i + 1]; ++t) {
                    const int b = s.lo_blk[t];
                    const int j = p->col[b];
                    /* L_ij = A_ij * Dinv_j   (:231) */
                    blk_gemm(n, blocks + (size_t)b * nn, Dinv + (size_t)j * nn, L + (size_t)b * nn);
                    /* D_i = D_i - L_ij * A_ji   (:234-239) */
                    blk_gemm(n, L + (size_t)b * nn, blocks + (size_t)s.lo_tr[t] * nn, tmp);
                    for (size_t k = 0; k < nn; ++k) Di[k] -= tmp[k];
                }
                /* Dinv_i = inv(D_i)   (:246-247) */
                int e = orc_inv(n, Di, inv);
                if (e) {
#pragma omp critical
                    { if (bad < 0 || i < bad) { bad = i; err = e; } }
                } else {
                    memcpy(Di, inv, sizeof(double) * nn);
                }
            }
            free(tmp); free(inv);
        }
    }
    if (bad_row) *bad_row = bad;
    struct_free(&s);
    return err;
}

/* ------------------------------------------------------------------------- */
/* ILU0 apply (precon_ILU0.f90:286-411)                                       */
/* ------------------------------------------------------------------------- */
int orc_ilu0_apply(const orc_pattern *p, const double *blocks, const double *L, const double *Dinv,
                   const double *v, double *z)
{
    const int n = p->n;
    const size_t nn = (size_t)n * n;
    orc_struct s;
    int err = struct_build(p, &s, 0);
    if (err) return err;
    /* z = v  (:305) */
    memcpy(z, v, sizeof(double) * (size_t)p->nrows * n);
    /* forward: z_i -= L_ij z_j over lower blocks, ascending rows (:336-359) */
    for (int l = 0; l < s.nlevels; ++l) {
#pragma omp parallel num_threads(orc_g_threads)
        {
            double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(static)
            for (int q = s.lev_ptr[l]; q < s.lev_ptr[l + 1]; ++q) {
                const int i = s.lev_rows[q];
                for (int t = s.lo_ptr[i]; t < s.lo_ptr[i + 1]; ++t) {
                    const int b = s.lo_blk[t];
                    blk_gemv_sub(n, L + (size_t)b * nn, z + (size_t)p->col[b] * n, z + (size_t)i * n, tmp);
                }
            }
            free(tmp);
        }
    }
    /* backward: z_i -= A_ik z_k over upper blocks; z_i = Dinv_i z_i (:363-398) */
    for (int l = 0; l < s.nblevels; ++l) {
#pragma omp parallel num_threads(orc_g_threads)
        {
            double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
            double *zi = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(static)
            for (int q = s.blev_ptr[l]; q < s.blev_ptr[l + 1]; ++q) {
                const int i = s.blev_rows[q];
                for (int t = s.up_ptr[i]; t < s.up_ptr[i + 1]; ++t) {
                    const int b = s.up_blk[t];
                    blk_gemv_sub(n, blocks + (size_t)b * nn, z + (size_t)p->col[b] * n, z + (size_t)i * n, tmp);
                }
                for (int r = 0; r < n; ++r) zi[r] = 0.0;
                blk_gemv_acc(n, Dinv + (size_t)i * nn, z + (size_t)i * n, zi, tmp);
                memcpy(z + (size_t)i * n, zi, sizeof(double) * (size_t)n);
            }
            free(tmp); free(zi);
        }
    }
    struct_free(&s);
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* Vector helpers (operator_chidg_dot.f90:30-89, type_chidg_vector.f90:1237-1300) */
/* ------------------------------------------------------------------------- */
double orc_dot(int64_t len, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for num_threads(orc_g_threads) reduction(+ : s) schedule(static)
    for (int64_t k = 0; k < len; ++k) s += a[k] * b[k];
    return s;
}

double orc_norm(int64_t len, const double *x) { return sqrt(orc_dot(len, x, x)); }

/* ------------------------------------------------------------------------- */
/* FGMRES (type_fgmres.f90:81-310)                                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    const orc_pattern *p;
    const double *blocks;
    int precon;
    const double *L, *Dinv;
    int64_t len;
} orc_sys;

static void sys_apply_precon(const orc_sys *S, const double *v, double *z)
{
    if (S->precon == ORC_PREC_ILU0) orc_ilu0_apply(S->p, S->blocks, S->L, S->Dinv, v, z);
    else memcpy(z, v, sizeof(double) * (size_t)S->len);   /* precon_identity */
}

/* CGS (:331-396): htmp_i = w . v_i for all i first, then w -= htmp_i v_i in
 * order, then h(1:j,j) += htmp.  MGS (:417-442): dot and update one at a time. */
static void orthogonalize(int orth, int64_t len, double *w, double **v, int j, double *hcol, double *htmp)
{
    if (orth == ORC_MGS) {
        for (int i = 0; i < j; ++i) {
            htmp[i] = orc_dot(len, w, v[i]);
            const double h = htmp[i];
            const double *vi = v[i];
#pragma omp parallel for num_threads(orc_g_threads) schedule(static)
            for (int64_t k = 0; k < len; ++k) w[k] = w[k] - h * vi[k];
        }
    } else {
        for (int i = 0; i < j; ++i) htmp[i] = orc_dot(len, w, v[i]);
        for (int i = 0; i < j; ++i) {
            const double h = htmp[i];
            const double *vi = v[i];
#pragma omp parallel for num_threads(orc_g_threads) schedule(static)
            for (int64_t k = 0; k < len; ++k) w[k] = w[k] - h * vi[k];
        }
    }
    for (int i = 0; i < j; ++i) hcol[i] += htmp[i];
}

static int fgmres_core(const orc_sys *S, const double *b, double *x, int nkrylov, double tol, double rtol,
                       int nmax, int orth, const orc_fgmres_params *inner, orc_fgmres_stats *st);

static int fgmres_core(const orc_sys *S, const double *b, double *x, int m, double tol, double rtol,
                       int nmax, int orth, const orc_fgmres_params *inner, orc_fgmres_stats *st)
{
    const int64_t len = S->len;
    const size_t vb = sizeof(double) * (size_t)len;
    if (m < 1) return ORC_ERR_ARG;
    double **v = (double **)calloc((size_t)m + 1, sizeof(double *));
    double **z = (double **)calloc((size_t)m + 1, sizeof(double *));
    double *h = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    double *p = (double *)calloc((size_t)m + 1, sizeof(double));
    double *c = (double *)calloc((size_t)m + 1, sizeof(double));
    double *s = (double *)calloc((size_t)m + 1, sizeof(double));
    double *htmp = (double *)calloc((size_t)m + 1, sizeof(double));
    double *y = (double *)calloc((size_t)m + 1, sizeof(double));
    double *x0 = (double *)malloc(vb), *r0 = (double *)malloc(vb), *w = (double *)malloc(vb);
    double *zr = inner ? (double *)malloc(vb) : NULL, *dz = inner ? (double *)malloc(vb) : NULL;
    double *hk = (double *)malloc(sizeof(double) * (size_t)m * m);
    double *hinv = (double *)malloc(sizeof(double) * (size_t)m * m);
    int err = ORC_OK;
    if (!v || !z || !h || !p || !c || !s || !htmp || !y || !x0 || !r0 || !w || !hk || !hinv ||
        (inner && (!zr || !dz))) { err = ORC_ERR_ALLOC; goto done; }
    for (int i = 0; i <= m; ++i) {
        v[i] = (double *)malloc(vb); z[i] = (double *)malloc(vb);
        if (!v[i] || !z[i]) { err = ORC_ERR_ALLOC; goto done; }
    }
#define H(i, j) h[(size_t)(j) * (m + 1) + (i)]
    /* x0 = 0, x = 0  (:132-136) */
    memset(x0, 0, vb);
    memset(x, 0, vb);
    st->niter = 0; st->nrestart = 0; st->nreorth = 0;
    double res = HUGE_VAL, r0norm_initial = 0.0;
    int conv_abs = 0, conv_rel = 0, conv_nmax = 0;
    while (res > tol) {
        memset(h, 0, sizeof(double) * (size_t)(m + 1) * m);
        memset(p, 0, sizeof(double) * (size_t)(m + 1));
        memset(c, 0, sizeof(double) * (size_t)(m + 1));
        memset(s, 0, sizeof(double) * (size_t)(m + 1));
        /* r0 = b - A x0 ; v1 = r0/||r0|| ; p1 = ||r0||  (:160-163) */
        orc_mv(S->p, S->blocks, x0, r0);
        for (int64_t k = 0; k < len; ++k) r0[k] = b[k] - r0[k];
        const double r0norm = orc_norm(len, r0);
        if (st->niter == 0) r0norm_initial = r0norm;
        if (r0norm == 0.0) {
            /* zero residual: x0 is exact.  The reference divides by zero here
             * (type_fgmres.f90:162); this restatement returns x = x0. */
            memcpy(x, x0, vb);
            res = 0.0;
            break;
        }
        for (int64_t k = 0; k < len; ++k) v[0][k] = r0[k] / r0norm;
        p[0] = r0norm;
        int nvecs = 0;
        for (int j = 0; j < m; ++j) {
            nvecs++;
            /* z_j = M v_j  (:175) */
            sys_apply_precon(S, v[j], z[j]);
            if (inner) {
                /* zr = v_j - A z_j ; inner solve ; z_j += dz  (:179-183) */
                orc_mv(S->p, S->blocks, z[j], zr);
                for (int64_t k = 0; k < len; ++k) zr[k] = v[j][k] - zr[k];
                orc_fgmres_stats ist;
                int e = fgmres_core(S, zr, dz, inner->inner_nkrylov, inner->inner_tol, inner->inner_rtol,
                                    inner->inner_nmax, inner->inner_orthogonalization, NULL, &ist);
                if (e) { err = e; goto done; }
                for (int64_t k = 0; k < len; ++k) z[j][k] = z[j][k] + dz[k];
            }
            /* w = A z_j ; nb = ||w||  (:189-191) */
            orc_mv(S->p, S->blocks, z[j], w);
            const double norm_before = orc_norm(len, w);
            orthogonalize(orth, len, w, v, j + 1, &H(0, j), htmp);
            /* selective re-orthogonalization, Giraud-Langou L_crit = 1 (:212-223) */
            double crit = 0.0;
            for (int i = 0; i <= j; ++i) crit += fabs(H(i, j));
            crit /= norm_before;
            if (crit >= 1.0) {
                orthogonalize(orth, len, w, v, j + 1, &H(0, j), htmp);
                st->nreorth++;
            }
            /* h(j+1,j) = ||w|| ; v_{j+1} = w / h(j+1,j)  (:229-231) */
            H(j + 1, j) = orc_norm(len, w);
            for (int64_t k = 0; k < len; ++k) v[j + 1][k] = w[k] / H(j + 1, j);
            /* previous Givens rotations on column j (:235-245) */
            for (int i = 0; i < j; ++i) {
                const double hij = c[i] * H(i, j) + s[i] * H(i + 1, j);
                const double hipj = -s[i] * H(i, j) + c[i] * H(i + 1, j);
                H(i, j) = hij;
                H(i + 1, j) = hipj;
            }
            /* new rotation (:249-259) */
            const double gam = hypot(H(j, j), H(j + 1, j));
            c[j] = H(j, j) / gam;
            s[j] = H(j + 1, j) / gam;
            H(j, j) = gam;
            H(j + 1, j) = 0.0;
            const double pj = c[j] * p[j];
            const double pjp = -s[j] * p[j];
            p[j] = pj;
            p[j + 1] = pjp;
            st->niter++;
            res = fabs(p[j + 1]);
            conv_abs = res < tol;
            conv_rel = (res / r0norm_initial) < rtol;
            conv_nmax = (st->niter >= nmax) && (nmax > 0);
            if (conv_abs || conv_rel || conv_nmax) break;
        }
        /* y = inv(H(1:k,1:k)) p(1:k)  (:287-288) */
        for (int jj = 0; jj < nvecs; ++jj)
            for (int ii = 0; ii < nvecs; ++ii) hk[(size_t)jj * nvecs + ii] = H(ii, jj);
        {
            int e = orc_inv(nvecs, hk, hinv);
            if (e) { err = e; goto done; }
        }
        for (int ii = 0; ii < nvecs; ++ii) {
            double acc = 0.0;
            for (int jj = 0; jj < nvecs; ++jj) acc += hinv[(size_t)jj * nvecs + ii] * p[jj];
            y[ii] = acc;
        }
        /* x = x0 + sum y_i z_i  (:292-295) */
        memcpy(x, x0, vb);
        for (int i = 0; i < nvecs; ++i) {
            const double yi = y[i];
            for (int64_t k = 0; k < len; ++k) x[k] = x[k] + yi * z[i][k];
        }
        if (conv_abs || conv_rel || conv_nmax) break;
        memcpy(x0, x, vb);
        st->nrestart++;
    }
#undef H
    /* residual_norm = ||b - A x||  (:305,515) */
    orc_mv(S->p, S->blocks, x, w);
    for (int64_t k = 0; k < len; ++k) w[k] = b[k] - w[k];
    st->residual_norm = orc_norm(len, w);
    for (int64_t k = 0; k < len; ++k)
        if (!isfinite(x[k])) { err = ORC_ERR_NAN; break; }
done:
    if (v) for (int i = 0; i <= m; ++i) free(v[i]);
    if (z) for (int i = 0; i <= m; ++i) free(z[i]);
    free(v); free(z); free(h); free(p); free(c); free(s); free(htmp); free(y);
    free(x0); free(r0); free(w); free(zr); free(dz); free(hk); free(hinv);
    return err;
}

int orc_fgmres_solve(const orc_pattern *p, const double *blocks, int precon, const double *L,
                     const double *Dinv, const double *b, double *x, const orc_fgmres_params *prm,
                     orc_fgmres_stats *st)
{
    if (!p || !prm || !st) return ORC_ERR_ARG;
    orc_sys S = {p, blocks, precon, L, Dinv, (int64_t)p->nrows * p->n};
    return fgmres_core(&S, b, x, prm->nkrylov, prm->tol, prm->rtol, prm->nmax, prm->orthogonalization,
                       prm->inner_fgmres ? prm : NULL, st);
}

