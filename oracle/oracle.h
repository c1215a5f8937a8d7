/*
 * oracle.h -- CPU FP64 restatement of ChiDG's Newton-Krylov linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the correctness checker and the
 * CPU baseline; it is never linked into, called by, or used as a fallback
 * for the product library (chidg_amd/csrc).  Only tests/, the smoke check
 * and the bench's cpu_baseline leg load it.
 *
 * Pinning (each function is checked against something it does not share
 * code with; see tests/test_oracle.py):
 *   orc_mv            reference known answers (chi_blks test), dense GEMV
 *   orc_inv           reference 3x3 known answer, numpy.linalg.inv, A*inv(A)=I
 *   orc_ilu0_factor   reference 3x3 known answer, dense block-ILU(0) (numpy),
 *                     1-D chain exactness (ILU0 == block LU)
 *   orc_ilu0_apply    reference 3x3 known answer, dense triangular solves
 *   orc_fgmres_solve  reference 3x3 / 2x2 known answers, numpy.linalg.solve,
 *                     1-iteration convergence when M = A^-1
 *   orc_fill_synthetic  statistical moments + bitwise equality with the
 *                     device generator (GPU test)
 *
 * Data model (mirrors the reference containers, flattened):
 *   rows      one block row per (element, time level), natural element order
 *             (type_domain_matrix.f90:253-305, mod_plot3d_utilities.f90:140-173)
 *   kind 0    lblks entry: interior face or diagonal coupling; takes part in ILU0
 *   kind 1    coupling-only entry (chi_blks / bc_blks / hb_blks): matvec only
 *             (operator_chidg_mv.f90:247-374; ILU0 never reads them,
 *             precon_ILU0.f90:213-243,336-389)
 *   blocks    dense n x n, column-major, n = nterms*nfields
 *             (type_densematrix.f90:60,177)
 */
#ifndef CHIDG_ORACLE_H
#define CHIDG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_ERR_ARG = 1,
    ORC_ERR_SINGULAR = 2,      /* mod_inv.f90:126-128 (FATAL there)            */
    ORC_ERR_NO_TRANSPOSE = 3,  /* type_domain_matrix.f90:620 (FATAL there)     */
    ORC_ERR_ALLOC = 4,
    ORC_ERR_NAN = 5
};

enum { ORC_PREC_IDENTITY = 0, ORC_PREC_ILU0 = 1 };
enum { ORC_CGS = 0, ORC_MGS = 1 };

typedef struct {
    int nrows;
    int n;
    int nnzb;
    const int32_t *row_ptr;   /* nrows+1                                    */
    const int32_t *col;       /* parent row of each block                   */
    const int32_t *kind;      /* 0 interior/diag, 1 coupling-only           */
    const int32_t *part;      /* rank of each row, or NULL (one rank): ILU0
                                 uses a kind-0 block only when its parent row
                                 has the row's rank (parent_proc == IRANK,
                                 precon_ILU0.f90:180,222,344,368)          */
} orc_pattern;

typedef struct {
    int    nkrylov;
    double tol;
    double rtol;
    int    nmax;
    int    orthogonalization;   /* ORC_CGS / ORC_MGS */
    int    inner_fgmres;
    int    inner_nkrylov;
    double inner_tol;
    double inner_rtol;
    int    inner_nmax;
    int    inner_orthogonalization;
} orc_fgmres_params;

typedef struct {
    int    niter;
    double residual_norm;   /* ||b - A x|| after the solve (type_fgmres.f90:305,515) */
    int    nrestart;
    int    nreorth;
} orc_fgmres_stats;

/* Number of OpenMP threads used by the *_par entry points (1 = serial). */
void orc_set_threads(int nthreads);
int  orc_get_max_threads(void);

/* y = A x over all blocks of every row (operator_chidg_mv.f90:200-374). */
int orc_mv(const orc_pattern *p, const double *blocks, const double *x, double *y);

/* Ainv = inv(A), LU with partial pivoting then inverse from the LU factors,
 * following LAPACK's published unblocked DGETF2 / DTRTI2 / DGETRI algorithm
 * (mod_inv.f90:99-150).  Returns ORC_ERR_SINGULAR on a zero pivot. */
int orc_inv(int n, const double *A, double *Ainv);

/* Natural-order block ILU(0) (precon_ILU0.f90:111-265).
 *   L     nnzb*n*n doubles; on return L[b] = A_ij * Dinv_j for every lower
 *         kind-0 block b; other slots are left untouched.
 *   Dinv  nrows*n*n doubles: inverted, updated diagonal blocks.
 * *bad_row receives the failing row on ORC_ERR_SINGULAR (may be NULL). */
int orc_ilu0_factor(const orc_pattern *p, const double *blocks, double *L, double *Dinv,
                    int *bad_row);

/* z = (LU)^-1 v with U taken from A's upper kind-0 blocks
 * (precon_ILU0.f90:286-411). */
int orc_ilu0_apply(const orc_pattern *p, const double *blocks, const double *L,
                   const double *Dinv, const double *v, double *z);

/* Restarted flexible GMRES (type_fgmres.f90:81-310): zero initial guess,
 * CGS or MGS with Giraud-Langou selective re-orthogonalization (L_crit = 1),
 * Givens rotations, restart, optional inner FGMRES correction.
 * precon = ORC_PREC_ILU0 uses (L, Dinv) from orc_ilu0_factor. */
int orc_fgmres_solve(const orc_pattern *p, const double *blocks, int precon,
                     const double *L, const double *Dinv, const double *b, double *x,
                     const orc_fgmres_params *prm, orc_fgmres_stats *st);

double orc_norm(int64_t len, const double *x);
double orc_dot(int64_t len, const double *a, const double *b);

/* Synthetic Jacobian values for a given pattern (bench/test input, not a
 * reference algorithm).  Entry (r,c) of block b:
 *     g   = approx N(0,1) from a counter hash of (seed, b, c*n+r)
 *     diag kind-0 block (col == row):  g*s_diag, plus 2.0 when r == c
 *     every other block:               g*s_off
 * Bitwise identical to the device generator (cdg_mat_fill_synthetic). */
void orc_fill_synthetic(const orc_pattern *p, uint64_t seed, double s_diag, double s_off,
                        double *blocks);
/* Same hash, one value per vector entry: g(seed, index) * scale. */
void orc_fill_vector(int64_t len, uint64_t seed, double scale, double *v);
double orc_synth_value(uint64_t seed, int64_t blk, int64_t elem);

#ifdef __cplusplus
}
#endif

#endif
