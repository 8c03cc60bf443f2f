/*
 * gcge_oracle.h -- CPU restatement of the GCGE generalized-CG eigensolver
 * (pase2017/GCGE-1.0, default code path) together with its CSR backend.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker for the HIP backend:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in the product links or calls it.
 *
 * Every function follows one reference routine; the reference file:line is
 * cited at each definition in gcge_oracle.c.  The restatement keeps the
 * reference's per-vector operation order (sequential dot products, the
 * branchy y = a*x + b*y, pointer-swap multivector semantics) so that, linked
 * against the same LAPACK, it reproduces the reference's iteration
 * trajectory bit for bit.  It is pinned by tests against the real reference
 * built from its own sources (oracle/_ref), against closed-form spectra and
 * against dense LAPACK eigensolves.
 */
#ifndef GCGE_ORACLE_H
#define GCGE_ORACLE_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CSR matrix, int32 indices, FP64 values (app/csr/gcge_app_csr.h:27-40). */
typedef struct {
    int n;               /* rows == columns                     */
    int nnz;
    const int *rowptr;   /* n+1                                 */
    const int *col;      /* nnz                                 */
    const double *val;   /* nnz                                 */
} orc_csr;

/* Solver parameters: the default-path subset of GCGE_PARA
 * (src/gcge_para.h:41-103) with the defaults of src/gcge_para.c:25-99. */
typedef struct {
    int    nev;
    int    ev_max_it;
    int    block_size;          /* 0 -> nev (gcge_para.c:364-368)            */
    int    given_init_evec;     /* 0 -> rand() init, else X0 is used         */
    double ev_tol;
    const char *conv_type;      /* "R", "O", "A", or the extension "NS"      */
    const char *orth_type;      /* "B" (only B-orthogonalization is restated) */
    const char *w_orth_type;    /* "multi", "gs", "bgs", "cbgs" or "scbgs"   */
    const char *p_orth_type;    /* "multi", "gs", "bgs" or "scbgs"           */
    double conv_omega_norm;
    double orth_zero_tol;       /* DBL_EPSILON is added in setup             */
    double reorth_tol;
    double scbgs_reorth_tol;    /* DBL_EPSILON is added in setup             */
    double criterion_tol;
    int    max_reorth_time;
    int    max_direct_orth_length;
    double multi_tol_for_lock;
    int    cg_max_it;
    double cg_rate;
    int    cg_type;
    int    if_use_bcg;
    int    dirichlet_boundary;
    int    opt_rr_eig_partly;
    unsigned seed;              /* srand(seed) before rand() initialisation  */
    int    w_orth_block_size;   /* scbgs W block; 0 -> max(block_size/5, 1)  */
    int    scbgs_wself_max_reorth_time;
} orc_params;

typedef struct {
    int    num_iter;
    int    num_unlock;
    double sum_res;
    int    bcg_calls;
    int    bcg_iters_total;     /* sum over BCG calls of (niter-1)           */
    int    spmv_count;          /* matrix-vector products (A and B)          */
} orc_info;

/* Fill p with the reference defaults (gcge_para.c:25-99). */
void orc_default_params(orc_params *p);

/* Solve A x = lambda B x for the lowest nev pairs (GCGE_Solve,
 * gcge_eigsol.c:65-542).  B == NULL means B = I.  X0 (n x nev column-major)
 * is read when given_init_evec != 0.  eval[nev], evec[n x nev] are outputs,
 * res[nev] receives the last residuals.  trace, when non-NULL, receives the
 * per-iteration report in the format of GCGE_PrintIterationInfo with
 * print_eval = 1 (gcge_para.c:390-437).  Returns 0 on success. */
int orc_solve(const orc_csr *A, const orc_csr *B, const orc_params *p,
              const double *X0, double *eval, double *evec, double *res,
              orc_info *info, FILE *trace);

/* Same as orc_solve, trace written to the named file (NULL -> none). */
int orc_solve_file(const orc_csr *A, const orc_csr *B, const orc_params *p,
                   const double *X0, double *eval, double *evec, double *res,
                   orc_info *info, const char *trace_path);

/* Kernel-level restatements used to check individual device ops. */
void orc_csr_matvec(const orc_csr *A, const double *x, double *y);
void orc_csr_matmultivec(const orc_csr *A, const double *X, int ldx,
                         double *Y, int ldy, int k);
/* a[(j)*lda + i] = <V_i, W_j>, i < p, j < q, sequential dots
 * (GCGE_Default_MultiVecInnerProd "ns", gcge_ops.c:239-261). */
void orc_multivec_innerprod(const double *V, int ldv, int p,
                            const double *W, int ldw, int q,
                            double *a, int lda);
/* Y_j = sum_i a[j*lda + i] X_i via the reference's axpby chain
 * (GCGE_Default_MultiVecLinearComb, gcge_ops.c:163-193). */
void orc_multivec_lincomb(const double *X, int ldx, int p,
                          double *Y, int ldy, int q,
                          const double *a, int lda, int n);
/* y = a x + b y with the branches of CSR_VecAxpby (gcge_app_csr.c:330-365). */
void orc_vec_axpby(double a, const double *x, double b, double *y, int n);
double orc_vec_dot(const double *x, const double *y, int n);

/* One call of the block CG inner solver (GCGE_BCG, gcge_cg.c:137-302) on
 * A W = RHS for k columns; W (n x k, col-major) holds the initial guess on
 * entry and the result on exit; RHS is overwritten with the final residual.
 * iters_out (may be NULL) receives niter-1. */
void orc_bcg(const orc_csr *A, double *RHS, double *W, int n, int k,
             int cg_max_it, double cg_rate, int *iters_out);

#ifdef __cplusplus
}
#endif
#endif
